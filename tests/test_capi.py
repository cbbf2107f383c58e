"""The C ABI without a GPU: library loads, exports every declared entry point,
and its host-only file readers agree with the oracle's and the reference's."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cases as CS
import dem_mph_amd as D
import pyoracle as P
from dem_mph_amd import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dem_mph.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(dmph_[a-z_]+)\s*\(", txt)))


def test_network_library_exports_every_declared_symbol():
    """libdemmph_nn.so exports each function include/dem_mph_nn.h declares."""
    from dem_mph_amd import surrogate as SG
    txt = open(os.path.join(ROOT, "include", "dem_mph_nn.h")).read()
    names = sorted(set(re.findall(r"\b(dmph_nn_[a-z0-9_]+)\s*\(", txt)))
    assert len(names) >= 15
    lib = SG.nnlib()
    for n in names:
        assert hasattr(lib, n), n


def test_library_exports_every_declared_symbol():
    lib = D.lib()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(D.EXPORTED) == set(names)


def test_params_reader_matches_oracle(tmp_path):
    for name in CS.CASES:
        c = CS.make_case(name)
        data, _ = CS.write_inputs(c, str(tmp_path))
        a = D.Params.read(data)
        b = P.parse_data(data)
        assert bytes(a) == bytes(b), name   # identical layout and values


SYSU = "/root/reference/results/SYSU"


@pytest.mark.skipif(not os.path.isdir(SYSU), reason="reference inputs absent")
@pytest.mark.parametrize("case", ["2mm", "0.8mm"])
def test_params_reader_on_shipped_data_files(case):
    a = D.Params.read(os.path.join(SYSU, case, "dem.data"))
    b = P.parse_data(os.path.join(SYSU, case, "dem.data"))
    assert bytes(a) == bytes(b)
    assert a.dt == 1.0e-5 and a.radius_ratio_a == 3.0 and a.young_modulus[2] > 0


def test_grid_reader_matches_python_reader(tmp_path):
    c = CS.make_case("mix3d")
    _, grid = CS.write_inputs(c, str(tmp_path))
    a = D.read_grid(grid)
    b = S.read_grid(grid)
    assert a.n == b.n
    np.testing.assert_array_equal(a.type, b.type)
    np.testing.assert_array_equal(a.pos, b.pos)
    np.testing.assert_array_equal(a.vel, b.vel)
    np.testing.assert_array_equal(a.domain_min, b.domain_min)


def test_grid_reader_extra_columns_and_errors(tmp_path):
    sc = S.dam_break_c1(seed=3, rigid=True, scale=0.4)
    p = str(tmp_path / "r.grid")
    S.write_grid(sc, p)
    g = D.read_grid(p)
    np.testing.assert_array_equal(g.body, sc.body)
    np.testing.assert_allclose(g.density, sc.density, rtol=1e-6)
    # truncated file: an error code, not a crash (the reference terminates)
    lines = open(p).read().splitlines()
    open(str(tmp_path / "t.grid"), "w").write("\n".join(lines[:10]) + "\n")
    with pytest.raises(D.DmphError):
        D.read_grid(str(tmp_path / "t.grid"))
    with pytest.raises(D.DmphError):
        D.read_grid(str(tmp_path / "missing.grid"))


def test_create_reports_errors_without_crashing(tmp_path):
    """Invalid domain (fewer than 3 cells) or no device: a return code and a message."""
    sc = S.Scene(d=0.02, dim=3, domain_min=np.zeros(3), domain_max=np.full(3, 0.1), type=np.ones(1, np.int32),
                 pos=np.full((1, 3), 0.05), vel=np.zeros((1, 3)))
    with pytest.raises(D.DmphError) as e:
        D.Solver(D.Params.from_dict(S.params_c1()), sc)
    assert "cell" in str(e.value) or "device" in str(e.value)


def test_ctypes_mirrors_match_the_header(tmp_path):
    """The package's ctypes structures lay out every field where the C header
    puts it (offsets and sizes from gcc on include/dem_mph.h), so a field
    appended to a struct on one side only fails here rather than as a
    misread option on the GPU."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    mirrors = {"dmph_params": D.Params, "dmph_grid": D.Grid, "dmph_options": D.Options, "dmph_info": D.Info}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dem_mph.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append('printf("%s sizeof %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in mirrors.items():
        assert got[(cname, "sizeof")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
