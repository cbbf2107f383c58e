"""Write tests/golden/cli_golden.npz: the files the reference's own program
writes, for the byte-level test of the DEM-MPH executable.

Run in a container that has the reference sources (it builds and runs
oracle/_ref/ref{2,3}d, see oracle/Makefile):

    python tests/golden/make_cli_golden.py

For each case (tests/cases.py scene and parameters, with a short end time and
output intervals of a few steps) the reference harness runs in --files mode:
the reference's main loop with its writers (writeProfFile, writeVtkFile), its
file naming sprintf(format, iStep) and output cadence
(source/DEM-MPH.cpp:419-580).  Two runs per case:
  - "fresh":   from the case's .grid;
  - "restart": from the .prof the fresh run wrote at the restart step, read
               back as a .grid (the reference resumes this way: Time from
               line 1, iStep = (int)(Time / Dt)).
The fixture holds the input files, every output file's bytes and the
"@ ... Output Time" log lines.  It is data (inputs and expected outputs).
"""
from __future__ import annotations

import glob
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import cases as CS  # noqa: E402

PROF_FMT = "run%03d.prof"
VTK_FMT = "run%03d.vtk"
# case -> (end time in steps, prof interval in steps, vtk interval in steps, restart step)
RUNS = {"mix3d": (8, 4, 4, 4), "wall2d": (4, 2, 2, 2)}


def cli_case(name: str):
    """The tests/cases.py case with the end time and output intervals of the CLI test."""
    c = CS.make_case(name)
    steps, every_prof, every_vtk, restart = RUNS[name]
    dt = c.params["Dt"]
    t0 = c.scene.time
    c.params["EndTime"] = t0 + steps * dt
    c.params["OutputInterval"] = every_prof * dt
    c.params["VtkOutputInterval"] = every_vtk * dt
    return c, restart


def at_lines(log_path: str):
    with open(log_path) as f:
        return [ln.rstrip("\n") for ln in f if ln.startswith("@ ")]


def run_reference(dim: int, data: str, grid: str, cwd: str):
    exe = os.path.join(ROOT, "oracle", "_ref", "ref%dd" % dim)
    subprocess.run([exe, "--files", data, grid, PROF_FMT, VTK_FMT, os.path.join(cwd, "run.log")], cwd=cwd, check=True,
                   capture_output=True)
    files = {}
    for p in sorted(glob.glob(os.path.join(cwd, "run*.prof")) + glob.glob(os.path.join(cwd, "run*.vtk")) +
                    [os.path.join(cwd, "output.vtk")]):
        with open(p, "rb") as f:
            files[os.path.basename(p)] = f.read()
    return files, at_lines(os.path.join(cwd, "run.log"))


def main() -> None:
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True, capture_output=True)
    out = {}
    meta = {}
    for name in RUNS:
        c, restart = cli_case(name)
        with tempfile.TemporaryDirectory() as tmp:
            fresh = os.path.join(tmp, "fresh")
            os.makedirs(fresh)
            data, grid = CS.write_inputs(c, fresh)
            for p in (data, grid):
                with open(p, "rb") as f:
                    out["%s/in/%s" % (name, os.path.basename(p)[len(name) + 1:])] = np.frombuffer(f.read(), np.uint8)
            files, log = run_reference(c.dim, data, grid, fresh)
            meta[name] = {"fresh": {"files": sorted(files), "log": log}}
            for fn, b in files.items():
                out["%s/fresh/%s" % (name, fn)] = np.frombuffer(b, np.uint8)
            # restart from the fresh run's .prof at the restart step
            iprof = int(round(c.scene.time / c.params["Dt"])) + restart
            rgrid = os.path.join(fresh, PROF_FMT % iprof)
            again = os.path.join(tmp, "restart")
            os.makedirs(again)
            files, log = run_reference(c.dim, data, rgrid, again)
            meta[name]["restart"] = {"files": sorted(files), "log": log, "from": PROF_FMT % iprof}
            for fn, b in files.items():
                out["%s/restart/%s" % (name, fn)] = np.frombuffer(b, np.uint8)
    out["meta_json"] = np.array(json.dumps(meta))
    path = os.path.join(HERE, "cli_golden.npz")
    np.savez_compressed(path, **out)
    print("%s: %d arrays, %.0f KB" % (path, len(out), os.path.getsize(path) / 1024))
    print(json.dumps(meta, indent=1)[:3000])


if __name__ == "__main__":
    main()
