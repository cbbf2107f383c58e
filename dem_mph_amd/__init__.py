"""MI355X-native DEM-MPH explicit step: Python access to libdemmph.so.

The library (HIP kernels for gfx950 behind the C ABI of include/dem_mph.h) is
built in-tree by ``dem_mph_amd.build``.  This module mirrors that ABI with
ctypes: ``Params.read`` (the reference's .data reader), ``read_grid`` (its
.grid reader), and ``Solver`` (create / step / fields / prof / vtk), so a
Python caller drives the same entry points a C or C++ caller would.  There is
no CPU fallback: without the built library or a HIP device every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libdemmph.so")

NTYPE = 6
OK, E_ARG, E_IO, E_FORMAT, E_DOMAIN, E_DEVICE, E_STATE, E_ALLOC = 0, -1, -2, -3, -4, -5, -6, -7
PREC_FP64, PREC_FP32 = 0, 1
SORT_CANONICAL, SORT_REFERENCE = 0, 1
REBUILD_SKIN, REBUILD_EVERY_STEP = 0, 1
DEBRIS_PARTICLES, DEBRIS_RIGID = 0, 1
ORDER_ORIGINAL, ORDER_SLOT = 0, 1

FIELDS = {
    "pos": (0, 3), "vel": (1, 3), "force": (2, 3), "type": (3, -1), "id": (4, -1), "density_a": (5, 1),
    "gravity_center": (6, 3), "pressure_a": (7, 1), "vol_strain_p": (8, 1), "divergence_p": (9, 1),
    "pressure_p": (10, 1), "mass": (11, 1), "alpha": (12, 1), "drag_c": (13, 1), "kn": (14, 1), "etan": (15, 1),
    "kt": (16, 1), "etat": (17, 1), "neighbor_count": (18, -1), "solid_neighbor_count": (19, -1),
    "key": (20, -1), "virial_pressure": (21, 1), "body": (22, -1), "kappa": (23, 1),
}

EXPORTED = ("dmph_default_options", "dmph_params_read", "dmph_grid_read", "dmph_grid_free", "dmph_create",
            "dmph_create_batch", "dmph_step", "dmph_sync", "dmph_time", "dmph_get_field", "dmph_set_field",
            "dmph_get_bodies", "dmph_get_info", "dmph_count_nonfinite",
            "dmph_virial", "dmph_pairs", "dmph_write_prof", "dmph_write_vtk", "dmph_set_timing", "dmph_get_timing",
            "dmph_autotune", "dmph_export_state",
            "dmph_boid_generate", "dmph_boid_error", "dmph_grid_write",
            "dmph_last_error", "dmph_destroy")

TIMING_GROUPS = ("predict", "rebin", "sweep1", "sweep2", "accel", "dem", "move", "end")


class DmphError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("dem_dt", C.c_double), ("output_interval", C.c_double),
        ("vtk_output_interval", C.c_double), ("end_time", C.c_double),
        ("radius_ratio_a", C.c_double), ("radius_ratio_p", C.c_double), ("radius_ratio_v", C.c_double),
        ("density", C.c_double * NTYPE), ("bulk_modulus", C.c_double * NTYPE),
        ("bulk_viscosity", C.c_double * NTYPE), ("shear_viscosity", C.c_double * NTYPE),
        ("surface_tension", C.c_double * NTYPE), ("interaction_ratio", (C.c_double * NTYPE) * NTYPE),
        ("gravity", C.c_double * 3), ("young_modulus", C.c_double * NTYPE), ("poisson_ratio", C.c_double * NTYPE),
        ("restitution", C.c_double * NTYPE), ("friction", C.c_double * NTYPE),
        ("wall_center", (C.c_double * 3) * NTYPE), ("wall_velocity", (C.c_double * 3) * NTYPE),
        ("wall_omega", (C.c_double * 3) * NTYPE),
    ]

    @classmethod
    def read(cls, path: str) -> "Params":
        """The reference's readDataFile (source/DEM-MPH.cpp:621-679)."""
        p = cls()
        p.dt = 1.0e100
        p.dem_dt = 1.0e100
        rc = lib().dmph_params_read(path.encode(), C.byref(p))
        if rc < 0:
            raise DmphError("cannot read %s" % path)
        return p

    @classmethod
    def from_dict(cls, d: Dict[str, object]) -> "Params":
        """From a dict keyed like the .data file (see dem_mph_amd.scenes.DEFAULT_PARAMS)."""
        p = cls()
        p.dt, p.dem_dt = d["Dt"], d["DEMDt"]
        p.output_interval, p.vtk_output_interval, p.end_time = d["OutputInterval"], d["VtkOutputInterval"], d["EndTime"]
        p.radius_ratio_a, p.radius_ratio_p, p.radius_ratio_v = d["RadiusRatioA"], d["RadiusRatioP"], d["RadiusRatioV"]
        for k, f in (("Density", "density"), ("BulkModulus", "bulk_modulus"), ("BulkViscosity", "bulk_viscosity"),
                     ("ShearViscosity", "shear_viscosity"), ("SurfaceTension", "surface_tension"),
                     ("YoungModulus", "young_modulus"), ("PoissonRatio", "poisson_ratio"),
                     ("RestitutionCoefficient", "restitution"), ("ViscousFrictionCoefficient", "friction")):
            arr = getattr(p, f)
            for t in range(NTYPE):
                arr[t] = d[k][t]
        for a in range(NTYPE):
            for b in range(NTYPE):
                p.interaction_ratio[a][b] = d["InteractionRatio"][a][b]
        for a in range(3):
            p.gravity[a] = d["Gravity"][a]
        for w in (4, 5):
            c = d["Wall%d" % w]
            for a in range(3):
                p.wall_center[w][a] = c["Center"][a]
                p.wall_velocity[w][a] = c["Velocity"][a]
                p.wall_omega[w][a] = c["Omega"][a]
        return p


class Grid(C.Structure):
    _fields_ = [("time", C.c_double), ("n", C.c_int), ("spacing", C.c_double), ("domain_min", C.c_double * 3),
                ("domain_max", C.c_double * 3), ("type", C.POINTER(C.c_int)), ("pos", C.POINTER(C.c_double)),
                ("vel", C.POINTER(C.c_double)), ("body", C.POINTER(C.c_int)), ("density", C.POINTER(C.c_double))]


class Options(C.Structure):
    _fields_ = [("dim", C.c_int), ("precision", C.c_int), ("sort_mode", C.c_int), ("rebuild", C.c_int),
                ("debris", C.c_int), ("device", C.c_int), ("compensated", C.c_int), ("graph", C.c_int),
                ("lanes", C.c_int), ("rigid_precision", C.c_int), ("block", C.c_int), ("cell_scale", C.c_double),
                ("wall_lanes", C.c_int)]


class Info(C.Structure):
    _fields_ = [("n", C.c_int), ("dim", C.c_int), ("cells", C.c_int), ("cell_count", C.c_int * 3),
                ("nbodies", C.c_int), ("substeps", C.c_int), ("precision", C.c_int),
                ("fluid_begin", C.c_int), ("fluid_end", C.c_int), ("solid_begin", C.c_int), ("solid_end", C.c_int),
                ("wall_begin", C.c_int), ("wall_end", C.c_int), ("cell_width", C.c_double * 3),
                ("n0a", C.c_double), ("n0p", C.c_double), ("max_radius", C.c_double), ("rebuilds", C.c_int64),
                ("nscenes", C.c_int), ("lanes", C.c_int * 2), ("block", C.c_int * 2), ("cell_scale", C.c_double),
                ("wall_lanes", C.c_int)]


_lib = None


def share_hip_runtime() -> None:
    """Load PyTorch (when installed) before any library of this package.

    PyTorch ships its own HIP runtime under the same soname as /opt/rocm's
    (libamdhip64.so.7).  With torch loaded first, the libraries here bind to
    that copy and the process has one runtime, so device pointers pass freely
    between torch and the C ABI.  In the other order torch loads a second
    runtime later, which then reports no GPU."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib() -> C.CDLL:
    """Load libdemmph.so (fails loudly when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DmphError("libdemmph.so is not built; run `python -m dem_mph_amd.build`")
        share_hip_runtime()
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.dmph_default_options.argtypes = [C.POINTER(Options)]
        L.dmph_params_read.argtypes = [C.c_char_p, C.POINTER(Params)]
        L.dmph_grid_read.argtypes = [C.c_char_p, C.POINTER(Grid)]
        L.dmph_boid_generate.argtypes = [C.c_char_p, C.POINTER(Grid)]
        L.dmph_boid_error.restype = C.c_char_p
        L.dmph_grid_write.argtypes = [C.c_char_p, C.POINTER(Grid)]
        L.dmph_grid_free.argtypes = [C.POINTER(Grid)]
        L.dmph_create.argtypes = [C.POINTER(Params), C.POINTER(Options), C.POINTER(Grid), C.POINTER(vp)]
        L.dmph_step.argtypes = [vp, C.c_int]
        L.dmph_sync.argtypes = [vp]
        L.dmph_time.argtypes = [vp, dp]
        L.dmph_get_field.argtypes = [vp, C.c_int, C.c_int, vp]
        L.dmph_get_bodies.argtypes = [vp, dp, dp, dp, dp, dp]
        L.dmph_get_info.argtypes = [vp, C.POINTER(Info)]
        L.dmph_virial.argtypes = [vp]
        L.dmph_count_nonfinite.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
        L.dmph_pairs.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), ip]
        L.dmph_write_prof.argtypes = [vp, C.c_char_p]
        L.dmph_write_vtk.argtypes = [vp, C.c_char_p]
        L.dmph_set_timing.argtypes = [vp, C.c_int]
        L.dmph_get_timing.argtypes = [vp, C.c_int, dp, C.POINTER(C.c_int64)]
        L.dmph_last_error.argtypes = [vp]
        L.dmph_last_error.restype = C.c_char_p
        L.dmph_destroy.argtypes = [vp]
        for f in EXPORTED:
            if f not in ("dmph_default_options", "dmph_grid_free", "dmph_destroy", "dmph_last_error", "dmph_boid_error"):
                getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def default_options(**kw) -> Options:
    o = Options()
    lib().dmph_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _scene_of(g: "Grid"):
    from . import scenes
    n = g.n
    t = np.ctypeslib.as_array(g.type, (n,)).astype(np.int32).copy() if n else np.zeros(0, np.int32)
    pos = np.ctypeslib.as_array(g.pos, (n, 3)).copy() if n else np.zeros((0, 3))
    vel = np.ctypeslib.as_array(g.vel, (n, 3)).copy() if n else np.zeros((0, 3))
    body = np.ctypeslib.as_array(g.body, (n,)).astype(np.int32).copy() if n and bool(g.body) else None
    dens = np.ctypeslib.as_array(g.density, (n,)).copy() if n and bool(g.density) else None
    dmin, dmax = np.array(g.domain_min[:]), np.array(g.domain_max[:])
    return scenes.Scene(d=g.spacing, dim=3, domain_min=dmin, domain_max=dmax, type=t, pos=pos, vel=vel,
                        body=body, density=dens, time=g.time)


def generate_boid(path: str, grid_path: Optional[str] = None):
    """The reference's scene generator (generator/generator.cpp): a .boid file
    into a scenes.Scene, and into a .grid file in the generator's layout when
    `grid_path` is given.  Raises on the generator's parse errors (the shapes
    read before the error are what the generator would still write)."""
    g = Grid()
    rc = lib().dmph_boid_generate(path.encode(), C.byref(g))
    try:
        if rc != OK:
            raise DmphError("%s: %s (code %d)" % (path, lib().dmph_boid_error().decode(), rc))
        if grid_path is not None:
            w = lib().dmph_grid_write(grid_path.encode(), C.byref(g))
            if w != OK:
                raise DmphError("cannot write %s (code %d)" % (grid_path, w))
        return _scene_of(g)
    finally:
        lib().dmph_grid_free(C.byref(g))


def read_grid(path: str):
    """The reference's readGridFile (source/DEM-MPH.cpp:681-851) into a scenes.Scene."""
    g = Grid()
    rc = lib().dmph_grid_read(path.encode(), C.byref(g))
    if rc != OK:
        raise DmphError("cannot read %s (code %d)" % (path, rc))
    try:
        return _scene_of(g)
    finally:
        lib().dmph_grid_free(C.byref(g))


class Solver:
    """One DEM-MPH scene on one GPU, driven through the C ABI -- or, given a
    list of scenes that share spacing and domain, a batch advanced together
    (dmph_create_batch); fields then cover all scenes, scene after scene
    (`scene_offsets`).  FP32 options: `lanes` and `block` (see `autotune`),
    `cell_scale` (cell width over the reference's, 1 to 2; negative: tuned at
    creation over 1, 1.25 and 1.5, reported by `info().cell_scale`)."""

    def __init__(self, params: Params, scene, precision: int = PREC_FP64, sort_mode: int = SORT_CANONICAL,
                 rebuild: int = REBUILD_SKIN, rigid: bool = False, dim: Optional[int] = None, device: int = 0,
                 compensated: bool = True, graph: bool = True, lanes: int = 0, rigid_precision: int = PREC_FP64,
                 block: int = 0, cell_scale: float = 1.0, wall_lanes: int = 0):
        self._arrays = []
        scenes = list(scene) if isinstance(scene, (list, tuple)) else [scene]
        grids = (Grid * len(scenes))()
        for k, sc in enumerate(scenes):
            self._fill_grid(grids[k], sc)
        first = scenes[0]
        self.options = default_options(dim=dim if dim else (first.dim if first.dim in (2, 3) else 0),
                                       precision=precision, sort_mode=sort_mode, rebuild=rebuild,
                                       debris=DEBRIS_RIGID if rigid else DEBRIS_PARTICLES, device=device,
                                       compensated=1 if compensated else 0, graph=1 if graph else 0, lanes=lanes,
                                       rigid_precision=rigid_precision, block=block, cell_scale=cell_scale,
                                       wall_lanes=wall_lanes)
        self.params = params
        h = C.c_void_p()
        if len(scenes) == 1:
            rc = lib().dmph_create(C.byref(params), C.byref(self.options), C.byref(grids[0]), C.byref(h))
        else:
            rc = lib().dmph_create_batch(C.byref(params), C.byref(self.options), grids, len(scenes), C.byref(h))
        if rc != OK:
            raise DmphError("dmph_create failed (%d): %s" % (rc, lib().dmph_last_error(None).decode()))
        self.h = h
        self.n = sum(sc.n for sc in scenes)
        self.scene_offsets = np.cumsum([0] + [sc.n for sc in scenes])

    def _fill_grid(self, g: "Grid", scene) -> None:
        t = np.ascontiguousarray(scene.type, np.int32)
        pos = np.ascontiguousarray(scene.pos, np.float64)
        vel = np.ascontiguousarray(scene.vel, np.float64)
        g.time, g.n, g.spacing = scene.time, scene.n, scene.d
        for a in range(3):
            g.domain_min[a] = scene.domain_min[a]
            g.domain_max[a] = scene.domain_max[a]
        g.type = t.ctypes.data_as(C.POINTER(C.c_int))
        g.pos = pos.ctypes.data_as(C.POINTER(C.c_double))
        g.vel = vel.ctypes.data_as(C.POINTER(C.c_double))
        self._arrays += [t, pos, vel]
        if scene.body is not None:
            b = np.ascontiguousarray(scene.body, np.int32)
            g.body = b.ctypes.data_as(C.POINTER(C.c_int))
            self._arrays.append(b)
        if scene.density is not None:
            dn = np.ascontiguousarray(scene.density, np.float64)
            g.density = dn.ctypes.data_as(C.POINTER(C.c_double))
            self._arrays.append(dn)

    def _check(self, rc: int, what: str) -> None:
        if rc != OK:
            raise DmphError("%s failed (%d): %s" % (what, rc, lib().dmph_last_error(self.h).decode()))

    def step(self, k: int = 1) -> None:
        self._check(lib().dmph_step(self.h, k), "dmph_step")

    def sync(self) -> None:
        self._check(lib().dmph_sync(self.h), "dmph_sync")

    def time(self) -> float:
        t = C.c_double()
        self._check(lib().dmph_time(self.h, C.byref(t)), "dmph_time")
        return t.value

    def field(self, name: str, order: int = ORDER_ORIGINAL) -> np.ndarray:
        fid, comps = FIELDS[name]
        if comps < 0:
            out = np.zeros(self.n, np.int32)
        else:
            out = np.zeros((self.n, comps) if comps > 1 else self.n, np.float64)
        self._check(lib().dmph_get_field(self.h, fid, order, out.ctypes.data_as(C.c_void_p)), "dmph_get_field")
        return out

    def info(self) -> Info:
        i = Info()
        self._check(lib().dmph_get_info(self.h, C.byref(i)), "dmph_get_info")
        return i

    def export_state(self, pos_ptr: int, vel_ptr: int) -> None:
        """FP32 positions / velocities [n][3], original order, into device memory at the given addresses."""
        self._check(lib().dmph_export_state(self.h, C.c_void_p(pos_ptr), C.c_void_p(vel_ptr)), "dmph_export_state")

    def autotune(self, reps: int = 5) -> tuple:
        """Pick lanes per particle for the FP32 sweeps by timing; returns (sweep1, sweep2)."""
        lanes = (C.c_int * 2)()
        self._check(lib().dmph_autotune(self.h, reps, lanes), "dmph_autotune")
        return lanes[0], lanes[1]

    def set_field(self, name: str, values: np.ndarray, order: int = ORDER_ORIGINAL) -> None:
        """Overwrite "pos" or "vel" (n x 3, float64) from host memory."""
        fid, comps = FIELDS[name]
        v = np.ascontiguousarray(values, np.float64).reshape(self.n, 3)
        self._check(lib().dmph_set_field(self.h, fid, order, v.ctypes.data_as(C.c_void_p)), "dmph_set_field")

    def bodies(self) -> Dict[str, np.ndarray]:
        nb = self.info().nbodies
        x, v, l = (np.zeros((nb, 3)) for _ in range(3))
        q = np.zeros((nb, 4)); m = np.zeros(nb)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        if nb:
            self._check(lib().dmph_get_bodies(self.h, dp(x), dp(v), dp(q), dp(l), dp(m)), "dmph_get_bodies")
        return {"x": x, "v": v, "q": q, "l": l, "mass": m}

    def virial(self) -> None:
        self._check(lib().dmph_virial(self.h), "dmph_virial")

    def count_nonfinite(self, name: str) -> int:
        """NaN or infinite values of a field, counted on the device (dmph_count_nonfinite)."""
        c = C.c_int64()
        self._check(lib().dmph_count_nonfinite(self.h, FIELDS[name][0], C.byref(c)), "dmph_count_nonfinite")
        return int(c.value)

    def pairs(self):
        """(ptr[n+1], idx[nnz]) of the neighbour list at the last rebinning, in slot order."""
        nnz = C.c_int64()
        ptr = np.zeros(self.n + 1, np.int64)
        self._check(lib().dmph_pairs(self.h, C.byref(nnz), ptr.ctypes.data_as(C.POINTER(C.c_int64)), None), "dmph_pairs")
        idx = np.zeros(max(nnz.value, 1), np.int32)
        self._check(lib().dmph_pairs(self.h, C.byref(nnz), ptr.ctypes.data_as(C.POINTER(C.c_int64)),
                                     idx.ctypes.data_as(C.POINTER(C.c_int))), "dmph_pairs")
        return ptr, idx[: nnz.value]

    def set_timing(self, on: bool) -> None:
        """Per-group device timing with HIP events on the solver's stream (resets the totals)."""
        self._check(lib().dmph_set_timing(self.h, 1 if on else 0), "dmph_set_timing")

    def timing(self) -> Dict[str, tuple]:
        """{group: (total_ms, launches)} accumulated since set_timing(True)."""
        out = {}
        for k, name in enumerate(TIMING_GROUPS):
            ms = C.c_double(); n = C.c_int64()
            self._check(lib().dmph_get_timing(self.h, k, C.byref(ms), C.byref(n)), "dmph_get_timing")
            out[name] = (ms.value, n.value)
        return out

    def write_prof(self, path: str) -> None:
        self._check(lib().dmph_write_prof(self.h, path.encode()), "dmph_write_prof")

    def write_vtk(self, path: str) -> None:
        self._check(lib().dmph_write_vtk(self.h, path.encode()), "dmph_write_vtk")

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().dmph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
