"""Build libdemmph.so (HIP kernels for gfx950 + C++ host) and the DEM-MPH executable, in-tree.

    python -m dem_mph_amd.build          # or dem_mph_amd.build.build()

Objects go to dem_mph_amd/_build/, the library to dem_mph_amd/libdemmph.so
and the command-line driver to dem_mph_amd/DEM-MPH.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libdemmph.so")
NN_LIB = os.path.join(PKG, "libdemmph_nn.so")
CLI = os.path.join(PKG, "DEM-MPH")
GEN_CLI = os.path.join(PKG, "GeneratorForMph")
ARCH = os.environ.get("DMPH_OFFLOAD_ARCH", "gfx950")


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


# Packed FP32 instructions are kept out of generated code: some of their forms
# have given wrong results on these boxes.
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build(verbose: bool = False) -> str:
    hipcc = _hipcc()
    os.makedirs(BUILD, exist_ok=True)
    jobs = [
        # FP64 parity kernels: no fused multiply-add, so the association matches the reference's
        [hipcc, "-x", "hip", *COMMON, *DEVICE, "-ffp-contract=off", "-c", os.path.join(CSRC, "dmph_k64.hip"),
         "-o", os.path.join(BUILD, "dmph_k64.o")],
        [hipcc, "-x", "hip", *COMMON, *DEVICE, "-ffp-contract=fast", "-c", os.path.join(CSRC, "dmph_k32.hip"),
         "-o", os.path.join(BUILD, "dmph_k32.o")],
        [hipcc, "-x", "c++", *COMMON, "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c",
         os.path.join(CSRC, "dmph_host.cpp"), "-o", os.path.join(BUILD, "dmph_host.o")],
        [hipcc, "-x", "c++", *COMMON, "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c",
         os.path.join(CSRC, "dmph_io.cpp"), "-o", os.path.join(BUILD, "dmph_io.o")],
        # scene generator (.boid -> .grid): host code, the generator's arithmetic unfused
        [hipcc, "-x", "c++", *COMMON, "-ffp-contract=off", "-c", os.path.join(CSRC, "dmph_boid.cpp"),
         "-o", os.path.join(BUILD, "dmph_boid.o")],
        # surrogate network kernels (bf16 MFMA), a library of their own
        [hipcc, "-x", "hip", *COMMON, *DEVICE, "-ffp-contract=fast", "-c", os.path.join(CSRC, "nn", "dmph_nn.hip"),
         "-o", os.path.join(BUILD, "dmph_nn.o")],
    ]
    with ThreadPoolExecutor(max_workers=4) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr, file=sys.stderr)
    objs = [os.path.join(BUILD, o) for o in ("dmph_k64.o", "dmph_k32.o", "dmph_host.o", "dmph_io.o", "dmph_boid.o")]
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp, *objs])
    os.replace(tmp, LIB)
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", NN_LIB + ".tmp", os.path.join(BUILD, "dmph_nn.o")])
    os.replace(NN_LIB + ".tmp", NN_LIB)
    _run([hipcc, "-x", "c++", *COMMON, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", os.path.join(CSRC, "dmph_cli.cpp"), "-o", CLI + ".tmp",
          "-L" + PKG, "-ldemmph", "-Wl,-rpath,$ORIGIN"])
    os.replace(CLI + ".tmp", CLI)
    _run([hipcc, "-x", "c++", *COMMON, os.path.join(CSRC, "dmph_gen_cli.cpp"), "-o", GEN_CLI + ".tmp",
          "-L" + PKG, "-ldemmph", "-Wl,-rpath,$ORIGIN"])
    os.replace(GEN_CLI + ".tmp", GEN_CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
