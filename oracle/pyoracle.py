"""ctypes access to the FP64 oracle and the reference driver.

TEST INFRASTRUCTURE ONLY: imported by the test suite, the smoke check and the
benchmark's CPU-baseline leg, never by the product package.

- ``Oracle``: this package's C restatement of the reference step
  (oracle/dmph_oracle.c, built into oracle/liboracle.so).
- ``run_reference`` / ``read_dump``: the reference's own passes, compiled from
  its sources into oracle/_ref/ref{2,3}d by oracle/Makefile, dumping every
  per-particle array after chosen steps.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess
from typing import Dict, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

SORT_REFERENCE, SORT_CANONICAL = 0, 1
REBUILD_SKIN, REBUILD_EVERY_STEP = 0, 1


class OrcParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("dem_dt", C.c_double), ("output_interval", C.c_double),
        ("vtk_output_interval", C.c_double), ("end_time", C.c_double),
        ("ratio_a", C.c_double), ("ratio_p", C.c_double), ("ratio_v", C.c_double),
        ("density", C.c_double * 6), ("bulk_modulus", C.c_double * 6), ("bulk_viscosity", C.c_double * 6),
        ("shear_viscosity", C.c_double * 6), ("surface_tension", C.c_double * 6),
        ("interaction", (C.c_double * 6) * 6), ("gravity", C.c_double * 3),
        ("young", C.c_double * 6), ("poisson", C.c_double * 6), ("restitution", C.c_double * 6),
        ("friction", C.c_double * 6), ("wall_center", (C.c_double * 3) * 6),
        ("wall_velocity", (C.c_double * 3) * 6), ("wall_omega", (C.c_double * 3) * 6),
    ]


class OrcOpts(C.Structure):
    _fields_ = [("dim", C.c_int), ("sort_mode", C.c_int), ("rebuild", C.c_int), ("rigid", C.c_int),
                ("threads", C.c_int)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: build it with `make -C oracle oracle`")
        L = C.CDLL(LIB_PATH)
        L.orc_parse_data.argtypes = [C.c_char_p, C.POINTER(OrcParams)]
        L.orc_parse_data.restype = C.c_int
        dp = C.POINTER(C.c_double); ip = C.POINTER(C.c_int)
        L.orc_create.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcOpts), C.c_double, C.c_int, C.c_double,
                                 dp, dp, ip, dp, dp, ip, dp]
        L.orc_create.restype = C.c_void_p
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_virial.argtypes = [C.c_void_p]
        L.orc_field_info.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_int)]
        L.orc_field_info.restype = C.c_int64
        L.orc_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        L.orc_get.restype = C.c_int
        L.orc_destroy.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def parse_data(path: str) -> OrcParams:
    p = OrcParams()
    if lib().orc_parse_data(path.encode(), C.byref(p)) < 0:
        raise FileNotFoundError(path)
    return p


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Oracle:
    """One FP64 oracle state, created in file order and run in slot order."""

    def __init__(self, params: OrcParams, scene, sort_mode: int = SORT_REFERENCE,
                 rebuild: int = REBUILD_SKIN, rigid: bool = False, threads: int = 0):
        self._keep = []
        t = np.ascontiguousarray(scene.type, np.int32)
        pos = np.ascontiguousarray(scene.pos, np.float64)
        vel = np.ascontiguousarray(scene.vel, np.float64)
        dmin = np.ascontiguousarray(scene.domain_min, np.float64)
        dmax = np.ascontiguousarray(scene.domain_max, np.float64)
        body = None if scene.body is None else np.ascontiguousarray(scene.body, np.int32)
        dens = None if scene.density is None else np.ascontiguousarray(scene.density, np.float64)
        self._keep += [t, pos, vel, dmin, dmax, body, dens]
        self.opts = OrcOpts(scene.dim, sort_mode, rebuild, 1 if rigid else 0, threads)
        self.params = params
        h = lib().orc_create(C.byref(params), C.byref(self.opts), scene.time, scene.n, scene.d, _dp(dmin), _dp(dmax),
                             _ip(t), _dp(pos), _dp(vel), None if body is None else _ip(body),
                             None if dens is None else _dp(dens))
        if not h:
            raise ValueError("oracle rejected the scene (cell grid needs >= 3 cells per axis)")
        self.h = h
        self.n = scene.n

    def step(self, k: int = 1) -> None:
        lib().orc_step(self.h, k)

    def virial(self) -> None:
        lib().orc_virial(self.h)

    def get(self, name: str) -> np.ndarray:
        dt = C.c_int()
        cnt = lib().orc_field_info(self.h, name.encode(), C.byref(dt))
        if cnt < 0 or dt.value < 0:
            raise KeyError(name)
        out = np.zeros(max(cnt, 0), dtype={0: np.float64, 1: np.int32, 2: np.int64}[dt.value])
        lib().orc_get(self.h, name.encode(), out.ctypes.data_as(C.c_void_p))
        return out

    def field(self, name: str, original_order: bool = True) -> np.ndarray:
        """A per-particle field reshaped to [N, ...], in original particle order by default."""
        a = self.get(name)
        n = self.n
        if a.size == 3 * n:
            a = a.reshape(n, 3)
        elif a.size == 9 * n:
            a = a.reshape(n, 3, 3)
        if original_order:
            ids = self.get("id")
            out = np.empty_like(a)
            out[ids] = a
            return out
        return a

    def scalars(self) -> Dict[str, float]:
        v = self.get("scalars")
        keys = ["time", "n0a", "n0p", "swa", "swg", "swp", "swv", "r2g", "cofk", "max_radius", "ra", "rg", "rp",
                "rv", "volume", "spacing", "dt", "dem_dt", "substeps", "rebuilds"]
        return dict(zip(keys, v.tolist()))

    def ranges(self) -> Dict[str, int]:
        v = self.get("ranges")
        keys = ["n", "ncell", "fb", "fe", "sb", "se", "wb", "we", "nbodies"]
        return dict(zip(keys, v.tolist()))

    def close(self) -> None:
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------------------------------------
# reference driver
# ----------------------------------------------------------------------------------------------

def reference_available(dim: int = 3) -> bool:
    return os.path.exists(os.path.join(REF_DIR, "ref%dd" % dim))


def run_reference(dim: int, data: str, grid: str, prefix: str, nsteps: int, every: int = 0,
                  threads: Optional[int] = None, timing: bool = False) -> subprocess.CompletedProcess:
    exe = os.path.join(REF_DIR, "ref%dd" % dim)
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    args = [exe, data, grid, prefix, str(nsteps)]
    if every:
        args.append(str(every))
    if timing:
        args.append("--time")
    return subprocess.run(args, check=True, capture_output=True, text=True, env=env)


def read_dump(path: str) -> Dict[str, np.ndarray]:
    """Read a reference dump written by oracle/ref_harness.cpp."""
    out: Dict[str, np.ndarray] = {}
    with open(path, "rb") as f:
        assert f.read(8) == b"DMPHDUMP"
        while True:
            nm = f.read(48)
            if len(nm) < 48:
                break
            code, = struct.unpack("<i", f.read(4))
            cnt, = struct.unpack("<q", f.read(8))
            dt = {0: np.float64, 1: np.int32, 2: np.int64}[code]
            out[nm.rstrip(b"\0").decode()] = np.frombuffer(f.read(cnt * np.dtype(dt).itemsize), dtype=dt).copy()
    return out


REF_NAMES = {  # oracle field name -> reference dump name
    "id": "ParticleIndex", "type": "Property", "pos": "Position", "vel": "Velocity", "force": "Force",
    "mass": "Mass", "density_a": "DensityA", "gravity_center": "GravityCenter", "pressure_a": "PressureA",
    "vol_strain_p": "VolStrainP", "divergence_p": "DivergenceP", "pressure_p": "PressureP", "kappa": "Kappa",
    "lambda": "Lambda", "mu": "Mu", "alpha": "alpha", "drag_c": "DragC", "kn": "NormalSpringConstant",
    "etan": "NormalViscousDampingConstant", "kt": "TangentialSpringConstant",
    "etat": "TangentialViscousDampingConstant", "virial_pressure": "VirialPressureAtParticle",
    "virial_stress": "VirialStressAtParticle", "build_pos": "NeighborCalculatedPosition",
    "neighbor_count": "NeighborCount", "fluid_neighbor_count": "FluidNeighborCount",
    "solid_neighbor_count": "SolidNeighborCount", "neighbor_ptr": "NeighborPtr", "neighbor_ind": "NeighborInd",
    "key": "CellIndex",
}


def params_from_dict(d: Dict[str, object]) -> OrcParams:
    """OrcParams from a parameter dict with the .data key names (see dem_mph_amd.scenes)."""
    p = OrcParams()
    p.dt = d["Dt"]; p.dem_dt = d["DEMDt"]; p.output_interval = d["OutputInterval"]
    p.vtk_output_interval = d["VtkOutputInterval"]; p.end_time = d["EndTime"]
    p.ratio_a = d["RadiusRatioA"]; p.ratio_p = d["RadiusRatioP"]; p.ratio_v = d["RadiusRatioV"]
    for k, f in (("Density", "density"), ("BulkModulus", "bulk_modulus"), ("BulkViscosity", "bulk_viscosity"),
                 ("ShearViscosity", "shear_viscosity"), ("SurfaceTension", "surface_tension"),
                 ("YoungModulus", "young"), ("PoissonRatio", "poisson"), ("RestitutionCoefficient", "restitution"),
                 ("ViscousFrictionCoefficient", "friction")):
        arr = getattr(p, f)
        for t in range(6):
            arr[t] = d[k][t]
    for a in range(6):
        for b in range(6):
            p.interaction[a][b] = d["InteractionRatio"][a][b]
    for a in range(3):
        p.gravity[a] = d["Gravity"][a]
    for w in (4, 5):
        c = d["Wall%d" % w]
        for a in range(3):
            p.wall_center[w][a] = c["Center"][a]; p.wall_velocity[w][a] = c["Velocity"][a]
            p.wall_omega[w][a] = c["Omega"][a]
    return p
