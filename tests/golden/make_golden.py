"""Write the golden fixtures under tests/golden/ from the reference's own passes.

Run in a container that has the reference sources (it builds and runs
oracle/_ref/ref{2,3}d, see oracle/Makefile):

    python tests/golden/make_golden.py

For every case of tests/cases.py it writes ``<case>.npz`` holding
  - the inputs exactly as the reference parsed them (``.grid`` values after
    ``%e`` rounding, .data parameters as JSON),
  - the reference's arrays after initialisation (step 0) and after the last
    step: permutation, cell keys, neighbour CSR, and every per-particle field.
The fixtures are data (inputs and expected outputs), not reference code.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import cases as CS  # noqa: E402
import pyoracle as P  # noqa: E402
from dem_mph_amd import scenes as S  # noqa: E402

FIELDS = ["ParticleIndex", "Property", "Position", "Velocity", "Force", "DensityA", "GravityCenter", "PressureA",
          "VolStrainP", "DivergenceP", "PressureP", "Kappa", "Mass", "alpha", "DragC", "NormalSpringConstant",
          "NormalViscousDampingConstant", "TangentialSpringConstant", "TangentialViscousDampingConstant",
          "NeighborCount", "FluidNeighborCount", "SolidNeighborCount", "CellIndex", "VirialPressureAtParticle",
          "NeighborCalculatedPosition"]
STEP0_FIELDS = ["ParticleIndex", "CellIndex", "NeighborPtr", "NeighborInd", "NeighborCount", "SolidNeighborCount",
                "VolStrainP", "DensityA", "GravityCenter", "Position"]


def main() -> None:
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all"], check=True, capture_output=True)
    for name in CS.CASES:
        c = CS.make_case(name)
        with tempfile.TemporaryDirectory() as tmp:
            data, grid = CS.write_inputs(c, tmp)
            P.run_reference(c.dim, data, grid, os.path.join(tmp, name), c.steps, c.steps)
            sc = S.read_grid(grid)
            r0 = P.read_dump(os.path.join(tmp, "%s_0.bin" % name))
            rN = P.read_dump(os.path.join(tmp, "%s_%d.bin" % (name, c.steps)))
        out = {
            "params_json": np.array(json.dumps(c.params)),
            "meta": np.array([c.dim, c.steps, sc.n], np.int64),
            "time0": np.array(sc.time), "spacing": np.array(sc.d),
            "domain_min": sc.domain_min, "domain_max": sc.domain_max,
            "in_type": sc.type, "in_pos": sc.pos, "in_vel": sc.vel,
            "scalars": rN["scalars"], "iscalars": rN["iscalars"],
        }
        for f in STEP0_FIELDS:
            out["s0_" + f] = r0[f]
        for f in FIELDS:
            out["sN_" + f] = rN[f]
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("%s: N=%d steps=%d -> %s (%.0f KB)" % (name, sc.n, c.steps, path, os.path.getsize(path) / 1024))


if __name__ == "__main__":
    main()
