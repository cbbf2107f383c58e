"""Golden outputs of the reference scene generator for tests/golden/boid/*.boid.

Runs the reference's own generator (generator/*.cpp, built from the reference
sources by `make -C oracle ref` into oracle/_ref/GeneratorForMph) on every
fixture and records the particle count, byte count and sha256 of the .grid it
writes in tests/golden/boid_golden.json.  For sysu_2mm.boid (the reference's
results/SYSU/2mm/DEM.boid) it also checks that the output equals the DEM.grid
the reference ships next to it, when that file is present.

    make -C oracle ref && python tests/golden/make_boid_golden.py
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_GEN = os.path.join(ROOT, "oracle", "_ref", "GeneratorForMph")
SHIPPED = {"sysu_2mm": "/root/reference/results/SYSU/2mm/DEM.grid"}


def run_generator(exe: str, boid: str, timeout: float = 20.0) -> bytes:
    """`exe name` in a scratch directory holding name.boid; the bytes of name.grid."""
    with tempfile.TemporaryDirectory() as d:
        shutil.copy(boid, os.path.join(d, "scene.boid"))
        subprocess.run([exe, "scene"], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                       timeout=timeout, check=True)
        with open(os.path.join(d, "scene.grid"), "rb") as f:
            return f.read()


def main() -> int:
    if not os.path.exists(REF_GEN):
        print("build the reference generator first: make -C oracle ref", file=sys.stderr)
        return 1
    out = {}
    for name in sorted(os.listdir(os.path.join(HERE, "boid"))):
        if not name.endswith(".boid"):
            continue
        key = name[:-5]
        data = run_generator(REF_GEN, os.path.join(HERE, "boid", name))
        n = int(data.split(b"\n")[1].split()[0])
        out[key] = {"n": n, "bytes": len(data), "sha256": hashlib.sha256(data).hexdigest()}
        if key in SHIPPED and os.path.exists(SHIPPED[key]):
            with open(SHIPPED[key], "rb") as f:
                assert f.read() == data, "reference generator output differs from the shipped " + SHIPPED[key]
            out[key]["matches_shipped_grid"] = os.path.relpath(SHIPPED[key], "/root/reference")
        print(key, out[key])
    with open(os.path.join(HERE, "boid_golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
