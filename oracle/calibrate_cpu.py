"""CPU calibration (test infrastructure): the reference solver built from its
own sources (oracle/_ref/ref3d, see oracle/Makefile) against this package's
FP64 restatement (liboracle.so) on the same input files, wall clock around the
step loop only, same thread count.

    python oracle/calibrate_cpu.py [--threads 8] [--c1-steps 100] [--c2-steps 3]

Runs in a container that has the reference sources.  Both sides read the same
.data/.grid files (the .grid values after %e rounding) and run the reference's
particle order (its bitonic network) and 0.05 d skin schedule; the oracle is
also timed in the canonical order the GPU bench compares against.  Prints one
JSON line per workload.  The reference's own timer (clock(),
source/DEM-MPH.cpp:588-593) sums CPU time over threads, so the harness's
--time mode measures wall time with omp_get_wtime instead.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import pyoracle as P  # noqa: E402
from dem_mph_amd import scenes as S  # noqa: E402


def time_reference(data: str, grid: str, prefix: str, steps: int, threads: int) -> float:
    r = P.run_reference(3, data, grid, prefix, steps, threads=threads, timing=True)
    m = re.search(r"seconds=([0-9.eE+-]+)", r.stdout)
    return float(m.group(1))


def time_oracle(params, grid: str, steps: int, threads: int, sort_mode: int) -> float:
    sc = S.read_grid(grid)
    o = P.Oracle(P.params_from_dict(params), sc, sort_mode=sort_mode, threads=threads)
    t0 = time.perf_counter()
    o.step(steps)
    el = time.perf_counter() - t0
    o.close()
    return el


def calibrate(name: str, sc, params, steps: int, threads: int) -> dict:
    with tempfile.TemporaryDirectory() as tmp:
        data, grid = os.path.join(tmp, name + ".data"), os.path.join(tmp, name + ".grid")
        S.write_data(params, data)
        S.write_grid(sc, grid, extra_columns=False)
        t_ref = time_reference(data, grid, os.path.join(tmp, name), steps, threads)
        t_ref_order = time_oracle(params, grid, steps, threads, P.SORT_REFERENCE)
        t_canon = time_oracle(params, grid, steps, threads, P.SORT_CANONICAL)
    n = sc.n
    return {"workload": name, "particles": n, "steps": steps, "threads": threads,
            "reference_s_per_step": t_ref / steps, "oracle_reference_order_s_per_step": t_ref_order / steps,
            "oracle_canonical_s_per_step": t_canon / steps,
            "reference_particle_steps_per_s": n * steps / t_ref,
            "oracle_particle_steps_per_s": n * steps / t_canon,
            "oracle_over_reference_speed": t_ref / t_canon}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--c1-steps", type=int, default=100)
    ap.add_argument("--c2-steps", type=int, default=3)
    a = ap.parse_args()
    if not P.reference_available(3):
        print("oracle/_ref/ref3d is not built (make -C oracle ref)", file=sys.stderr)
        return 2
    # loose DEM grains: the reference's own debris model
    print(json.dumps(calibrate("dam_break_c1", S.dam_break_c1(seed=0), S.params_c1(), a.c1_steps, a.threads)),
          flush=True)
    print(json.dumps(calibrate("dam_break_c2", S.dam_break_c2(seed=1, rigid=False), S.params_c2(), a.c2_steps,
                               a.threads)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
