"""The DEM-MPH executable (dmph_cli.cpp on the HIP library) against the
reference's own program, file by file.

tests/golden/cli_golden.npz holds what the reference writes for two cases
(tests/golden/make_cli_golden.py): the output.vtk before the loop, the .prof
and .vtk files named by sprintf(format, iStep) at the reference's cadence
(source/DEM-MPH.cpp:419-580), the "@ ... Output Time" log lines, and the same
after a restart from one of its own .prof files read back as the .grid
(Time from line 1, iStep = (int)(Time / Dt)).  Here the executable runs the
same inputs in FP64 with the reference's particle order, and must produce the
same file names, the same log lines and the same bytes.  The only difference
allowed is a documented one: a number printed with %e may differ in its last
digit when the FP64 values differ in their last bits (device pow/log), at most
one unit in the 7th significant digit, and in at most 1 in 10^4 numbers.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

import cases as CS  # noqa: E402
import make_cli_golden as MG  # noqa: E402

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "dem_mph_amd", "DEM-MPH")
FIX = np.load(os.path.join(ROOT, "tests", "golden", "cli_golden.npz"))
META = json.loads(str(FIX["meta_json"]))


def _run_cli(data, grid, cwd):
    env = dict(os.environ, DMPH_PRECISION="fp64", DMPH_SORT="reference")
    r = subprocess.run([EXE, data, grid, MG.PROF_FMT, MG.VTK_FMT, os.path.join(cwd, "run.log")], cwd=cwd, env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    files = sorted(f for f in os.listdir(cwd) if f.endswith(".prof") or f.endswith(".vtk"))
    return files, MG.at_lines(os.path.join(cwd, "run.log"))


def _compare(ours: bytes, ref: bytes, what: str):
    """Equal bytes, or equal lines up to last-digit differences of %e numbers.
    Returns (numbers compared, numbers differing in the last digit)."""
    if ours == ref:
        return 0, 0
    a, b = ours.decode().split("\n"), ref.decode().split("\n")
    assert len(a) == len(b), "%s: %d lines, reference %d" % (what, len(a), len(b))
    nums = diff = 0
    for k, (x, y) in enumerate(zip(a, b)):
        if x == y:
            nums += len(y.split())
            continue
        tx, ty = x.split(" "), y.split(" ")
        assert len(tx) == len(ty), "%s line %d: %r vs %r" % (what, k + 1, x, y)
        for u, v in zip(tx, ty):
            nums += 1
            if u == v:
                continue
            mu, mv = u.split("e"), v.split("e")
            assert len(mu) == 2 and mu[1] == mv[1], "%s line %d: %r vs %r" % (what, k + 1, x, y)
            # same exponent: mantissas of 7 significant digits one unit apart at most
            assert abs(float(mu[0]) - float(mv[0])) <= 1.0000001e-6, "%s line %d: %r vs %r" % (what, k + 1, x, y)
            diff += 1
    return nums, diff


@pytest.mark.parametrize("name", sorted(MG.RUNS))
def test_cli_writes_the_reference_files(name, tmp_path):
    c, restart = MG.cli_case(name)
    fresh = tmp_path / "fresh"
    fresh.mkdir()
    data, grid = CS.write_inputs(c, str(fresh))
    # the inputs are the bytes the reference read
    for key, path in (("data", data), ("grid", grid)):
        with open(path, "rb") as f:
            assert f.read() == FIX["%s/in/%s" % (name, key)].tobytes(), key
    total = [0, 0]
    for run, grid_in, cwd in (("fresh", grid, fresh), ("restart", None, tmp_path / "restart")):
        if run == "restart":
            cwd.mkdir()
            grid_in = str(fresh / META[name]["restart"]["from"])
            # the restart reads our own .prof: it must be the reference's, byte for byte
            with open(grid_in, "rb") as f:
                assert f.read() == FIX["%s/fresh/%s" % (name, META[name]["restart"]["from"])].tobytes()
        print("CLI %s %s" % (name, run), flush=True)
        files, log = _run_cli(data, grid_in, str(cwd))
        assert files == META[name][run]["files"]
        assert log == META[name][run]["log"]
        for fn in files:
            with open(os.path.join(str(cwd), fn), "rb") as f:
                n, d = _compare(f.read(), FIX["%s/%s/%s" % (name, run, fn)].tobytes(), "%s/%s/%s" % (name, run, fn))
            total[0] += n
            total[1] += d
    sys.stdout.write("\n[cli-%s] numbers in differing files %d, last-digit differences %d\n" % (name, *total))
    assert total[1] <= max(1, total[0] // 10000)
