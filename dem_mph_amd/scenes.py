"""Seeded scene generation in the reference's particle input format.

The reference reads a ``.grid`` file written by its generator
(generator/generator.cpp:948-971): line 1 ``Time``, line 2
``N d xmin xmax ymin ymax zmin zmax``, then one ``type x y z vx vy vz`` line
per particle, every float printed with ``%e`` (7 significant digits).  The
solver parses it at source/DEM-MPH.cpp:681-851 with ``sscanf`` of seven fields,
so two optional trailing columns (``body_id``, ``density``) used by this
package's rigid-debris model stay readable by the reference.

Particle types follow source/DEM-MPH.cpp:46-52: 0-1 fluid, 2-3 solid (debris),
4-5 wall.

The scene builder here is this package's own: lattice blocks at spacing ``d``
(points at ``lo + (k + 1/2) d``), optional uniform jitter of ``+-jitter*d`` per
coordinate drawn from a splitmix64 stream, tanks walled by ``layers`` lattice
layers, and a periodic domain padded by at least ``MaxRadius + 0.1 d`` so that
periodic images never interact (the reference's domain is periodic on all
axes, source/DEM-MPH.cpp:3153-3164).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

FLUID, FLUID_B, SOLID, SOLID_B, WALL, WALL_B = range(6)

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """``n`` outputs of the splitmix64 generator (Steele, Lea, Flood 2014)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x632BE59BD9B4E019 + stream * 0x8CB92BA72F3D8DD7) & 0xFFFFFFFFFFFFFFFF)
        z = base + _GOLDEN * (np.arange(1, n + 1, dtype=np.uint64))
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int, stream: int = 0, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    """Uniform doubles in [lo, hi) from the top 53 bits of splitmix64."""
    u = (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def lattice(lo: Sequence[float], hi: Sequence[float], d: float, dim: int = 3) -> np.ndarray:
    """Lattice points ``lo + (k + 1/2) d`` filling the box [lo, hi)."""
    axes = []
    for a in range(3):
        if dim == 2 and a == 2:
            axes.append(np.array([lo[2] + 0.5 * d]))
            continue
        n = int(round((hi[a] - lo[a]) / d))
        axes.append(lo[a] + (np.arange(n) + 0.5) * d)
    g = np.meshgrid(*axes, indexing="ij")
    return np.stack([x.ravel() for x in g], axis=1)


@dataclasses.dataclass
class Scene:
    """Particles plus domain, in the reference's ``.grid`` terms."""
    d: float
    dim: int
    domain_min: np.ndarray
    domain_max: np.ndarray
    type: np.ndarray            # int32 [N]
    pos: np.ndarray             # float64 [N,3]
    vel: np.ndarray             # float64 [N,3]
    body: Optional[np.ndarray] = None      # int32 [N], -1 = not in a rigid body
    density: Optional[np.ndarray] = None   # float64 [N], per-particle density override
    time: float = 0.0

    @property
    def n(self) -> int:
        return int(self.type.shape[0])

    def counts(self) -> Dict[str, int]:
        t = self.type
        return {"fluid": int(np.sum(t < 2)), "solid": int(np.sum((t >= 2) & (t < 4))),
                "wall": int(np.sum(t >= 4)), "bodies": int(self.body.max() + 1) if self.body is not None and self.body.size else 0}

    def rounded(self) -> "Scene":
        """The scene as the reference reads it back: every float through ``%e``."""
        rt = lambda a: np.array([float("%e" % v) for v in np.ravel(a)]).reshape(np.shape(a))
        return dataclasses.replace(self, pos=rt(self.pos), vel=rt(self.vel),
                                   domain_min=rt(self.domain_min), domain_max=rt(self.domain_max),
                                   density=None if self.density is None else rt(self.density))


def write_grid(scene: Scene, path: str, extra_columns: bool = True) -> None:
    """Write ``scene`` in the generator's format (generator/generator.cpp:952-965).

    With ``extra_columns`` and a rigid-body assignment, each particle line
    carries two more fields, ``body_id density``, which the reference's seven
    field ``sscanf`` (source/DEM-MPH.cpp:808) ignores.
    """
    n = scene.n
    with open(path, "w") as f:
        f.write("%f\n" % scene.time)
        f.write("%d %e  %e %e %e  %e %e %e\n" % (
            n, scene.d, scene.domain_min[0], scene.domain_max[0], scene.domain_min[1],
            scene.domain_max[1], scene.domain_min[2], scene.domain_max[2]))
        p, v, t = scene.pos, scene.vel, scene.type
        with_body = extra_columns and scene.body is not None
        lines = []
        for i in range(n):
            s = "%d    %e %e %e  %e %e %e " % (t[i], p[i, 0], p[i, 1], p[i, 2], v[i, 0], v[i, 1], v[i, 2])
            if with_body:
                dens = scene.density[i] if scene.density is not None else -1.0
                s += " %d %e" % (scene.body[i], dens)
            lines.append(s)
        f.write("\n".join(lines))
        f.write("\n")


def read_grid(path: str) -> Scene:
    """Parse a ``.grid``/``.prof`` file (source/DEM-MPH.cpp:681-851), extra columns included."""
    with open(path) as f:
        time = float(f.readline().split()[0])
        hdr = f.readline().split()
        n = int(hdr[0]); d = float(hdr[1])
        vals = [float(x) for x in hdr[2:8]]
        dmin = np.array([vals[0], vals[2], vals[4]]); dmax = np.array([vals[1], vals[3], vals[5]])
        typ = np.zeros(n, np.int32); pos = np.zeros((n, 3)); vel = np.zeros((n, 3))
        body = np.full(n, -1, np.int32); dens = np.full(n, -1.0); have_extra = False
        for i in range(n):
            parts = f.readline().split()
            if not parts:
                break
            typ[i] = int(parts[0]); pos[i] = [float(x) for x in parts[1:4]]; vel[i] = [float(x) for x in parts[4:7]]
            if len(parts) >= 9:
                have_extra = True; body[i] = int(parts[7]); dens[i] = float(parts[8])
    dim = 2 if abs((dmax[2] - dmin[2]) - d) < 1e-12 * max(1.0, d) else 3
    return Scene(d=d, dim=dim, domain_min=dmin, domain_max=dmax, type=typ, pos=pos, vel=vel,
                 body=body if have_extra else None, density=dens if have_extra else None, time=time)


# ----------------------------------------------------------------------------------------------
# .data parameter files (keys of source/DEM-MPH.cpp:634-660)
# ----------------------------------------------------------------------------------------------

DEFAULT_PARAMS: Dict[str, object] = {
    # results/SYSU/2mm/dem.data, the only parameter set the reference ships
    "Dt": 1.0e-5, "DEMDt": 1.0e-5, "OutputInterval": 1.0, "VtkOutputInterval": 1.0e-2, "EndTime": 2.0e2,
    "RadiusRatioA": 3.0, "RadiusRatioP": 2.5, "RadiusRatioV": 3.0,
    "Density": [2.0e3, 1.0e3, 8.0e3, 1.0e3, 5.0e3, 5.0e3],
    "BulkModulus": [1e5, 1.0e5, 0.0, 1e4, 1.0e6, 1e4],
    "BulkViscosity": [1.0, 1.0e2, 0.0, 1e-1, 1.0e3, 1e-1],
    "ShearViscosity": [1.0e-2, 5.0e-4, 1e-3, 1e-1, 1e3, 1e3],
    "SurfaceTension": [0.072, 0.1, 0.0, 0.0, 0.0, 0.0],
    "InteractionRatio": [[0.0] * 6 for _ in range(6)],
    "Gravity": [0.0, -9.81, 0.0],
    "YoungModulus": [0.0, 0.0, 1.0e11, 1.0e10, 5.0e10, 6.0e10],
    "PoissonRatio": [0.0, 0.0, 3.7e-1, 2.0e-1, 4.0e-1, 1e-1],
    "RestitutionCoefficient": [0.0, 0.0, 0.90, 0.90, 0.60, 0.9],
    "ViscousFrictionCoefficient": [0.0, 0.0, 0.37, 0.4, 0.4, 0.1],
    "Wall4": {"Center": [0.0, 0.0, 0.0], "Velocity": [0.0, 0.0, 0.0], "Omega": [0.0, 0.0, 0.0]},
    "Wall5": {"Center": [0.0, 0.0, 0.0], "Velocity": [-1.0, 0.0, 0.0], "Omega": [0.0, 0.0, 0.0]},
}


def make_params(**over) -> Dict[str, object]:
    import copy
    p = copy.deepcopy(DEFAULT_PARAMS)
    for k, v in over.items():
        p[k] = copy.deepcopy(v)
    return p


def write_data(params: Dict[str, object], path: str) -> None:
    """Write a ``.data`` file the reference's ``readDataFile`` accepts."""
    six = lambda v: " ".join("%.17g" % x for x in v)
    with open(path, "w") as f:
        f.write("#######\n")
        for k in ("Dt", "DEMDt", "OutputInterval", "VtkOutputInterval", "EndTime",
                  "RadiusRatioA", "RadiusRatioP", "RadiusRatioV"):
            f.write("%s %.17g\n" % (k, params[k]))
        for k in ("Density", "BulkModulus", "BulkViscosity", "ShearViscosity", "SurfaceTension"):
            f.write("%s %s\n" % (k, six(params[k])))
        for t in range(6):
            f.write("InteractionRatio(Type%d) %s\n" % (t, six(params["InteractionRatio"][t])))
        f.write("Gravity %s\n" % " ".join("%.17g" % x for x in params["Gravity"]))
        f.write("##DEM Parameters##\n")
        for k in ("YoungModulus", "PoissonRatio", "RestitutionCoefficient", "ViscousFrictionCoefficient"):
            f.write("%s %s\n" % (k, six(params[k])))
        for w in ("Wall4", "Wall5"):
            c = params[w]
            f.write("%s  Center %s Velocity %s Omega %s\n" % (
                w, " ".join("%.17g" % x for x in c["Center"]), " ".join("%.17g" % x for x in c["Velocity"]),
                " ".join("%.17g" % x for x in c["Omega"])))


# ----------------------------------------------------------------------------------------------
# scene builders
# ----------------------------------------------------------------------------------------------

def _pad_domain(lo: np.ndarray, hi: np.ndarray, d: float, max_radius_ratio: float, dim: int):
    """Domain = particle bounding box padded by MaxRadius + 0.1 d + one spacing."""
    pad = (max_radius_ratio + 0.1) * d + d
    dmin = lo - pad; dmax = hi + pad
    if dim == 2:
        dmin[2] = 0.0; dmax[2] = d
    return dmin, dmax


def tank_walls(interior: Sequence[float], d: float, layers: int, closed_top: bool, dim: int) -> np.ndarray:
    """Wall lattice points in ``layers`` shells around the box [0, interior):
    the lattice of the padded box without its interior points, in lattice
    order (x slowest), built one x plane at a time so that large tanks never
    hold the full box lattice."""
    L = np.asarray(interior, float)
    lo = -layers * d * np.ones(3); hi = L + layers * d
    if not closed_top:
        up = 2 if dim == 3 else 1
        hi[up] = L[up]
    if dim == 2:
        lo[2] = 0.0; hi[2] = d
    full = lattice(np.array([lo[0], lo[1], lo[2]]), np.array([lo[0] + d, hi[1], hi[2]]), d, dim)   # one x plane
    nx = int(round((hi[0] - lo[0]) / d))
    inner_yz = np.ones(len(full), bool)
    for a in range(1, dim):
        inner_yz &= (full[:, a] > 0.0) & (full[:, a] < L[a])
    out = []
    for k in range(nx):
        x = lo[0] + (k + 0.5) * d
        pts = full.copy()
        pts[:, 0] = x
        out.append(pts if not (0.0 < x < L[0]) else pts[~inner_yz])
    return np.concatenate(out)


def dam_break(d: float, interior: Sequence[float], fluid: Sequence[float], layers: int = 3,
              debris: Sequence[Tuple[Sequence[float], Sequence[int], float]] = (),
              dim: int = 3, closed_top: bool = False, jitter: float = 0.01, seed: int = 0,
              max_radius_ratio: float = 3.0, rigid: bool = False,
              fluid_velocity: Sequence[float] = (0.0, 0.0, 0.0),
              debris_velocity: Sequence[float] = (0.0, 0.0, 0.0)) -> Scene:
    """Dam break in a walled tank.

    ``interior`` is the tank's inner box [0, interior); the fluid block sits in
    the corner [0, fluid); each debris entry is ``(origin, (nx, ny, nz), density)``
    — a cuboid of lattice particles whose first point is ``origin + d/2``.
    Gravity points along -z in 3D (``dim=3``) and -y in 2D, as in the
    reference's shipped 2D case (results/SYSU/2mm/dem.data:21).
    """
    fl = lattice(np.zeros(3), np.asarray(fluid, float), d, dim)
    walls = tank_walls(interior, d, layers, closed_top, dim)
    deb_pts: List[np.ndarray] = []; deb_body: List[np.ndarray] = []; deb_rho: List[np.ndarray] = []
    for b, (origin, cnt, rho) in enumerate(debris):
        hi = np.asarray(origin, float) + np.asarray(cnt, float) * d
        if dim == 2:
            hi[2] = d
        pts = lattice(np.asarray(origin, float), hi, d, dim)
        deb_pts.append(pts); deb_body.append(np.full(len(pts), b, np.int32)); deb_rho.append(np.full(len(pts), rho))
    sol = np.concatenate(deb_pts) if deb_pts else np.zeros((0, 3))
    pos = np.concatenate([fl, sol, walls])
    typ = np.concatenate([np.full(len(fl), FLUID, np.int32), np.full(len(sol), SOLID, np.int32),
                          np.full(len(walls), WALL, np.int32)])
    vel = np.zeros_like(pos)
    vel[: len(fl)] = np.asarray(fluid_velocity, float)
    vel[len(fl): len(fl) + len(sol)] = np.asarray(debris_velocity, float)
    if jitter > 0:
        j = uniform(seed, pos.size, stream=1, lo=-jitter * d, hi=jitter * d).reshape(pos.shape)
        if dim == 2:
            j[:, 2] = 0.0
        pos = pos + j
    lo = pos.min(axis=0); hi = pos.max(axis=0)
    dmin, dmax = _pad_domain(lo, hi, d, max_radius_ratio, dim)
    n_sol = len(sol)
    body = None; dens = None
    if rigid:
        body = np.full(len(pos), -1, np.int32); dens = np.full(len(pos), -1.0)
        if n_sol:
            body[len(fl): len(fl) + n_sol] = np.concatenate(deb_body)
            dens[len(fl): len(fl) + n_sol] = np.concatenate(deb_rho)
    return Scene(d=d, dim=dim, domain_min=dmin, domain_max=dmax, type=typ, pos=pos, vel=vel, body=body, density=dens)


# ----------------------------------------------------------------------------------------------
# the five workloads
# ----------------------------------------------------------------------------------------------

def dam_break_c1(seed: int = 0, rigid: bool = False, scale: float = 1.0) -> Scene:
    """3D dam break with debris, d = 0.02 m.

    Tank interior 0.8 x 0.4 x 0.6 m, three wall layers, open top; fluid block
    0.3 x 0.4 x 0.3 m (15 x 20 x 15 = 4,500 particles) in the x=0 corner; four
    4x4x4 debris cubes resting on the floor at x = 0.50-0.58 and 0.62-0.70 m;
    positions jittered by +-0.01 d.  ``scale`` < 1 shrinks every length (same
    d) for quick tests.
    """
    d = 0.02
    s = scale
    interior = (0.8 * s, 0.4 * s, 0.6 * s)
    fluid = (0.3 * s, 0.4 * s, 0.3 * s)
    deb = []
    xs = (0.50 * s, 0.62 * s) if s >= 1 else (interior[0] - 10 * d, interior[0] - 5 * d)
    ys = (0.08 * s, 0.24 * s) if s >= 1 else (0.0, interior[1] - 4 * d)
    for x in xs:
        for y in ys:
            deb.append(((x, y, 0.0), (4, 4, 4), 800.0))
    return dam_break(d, interior, fluid, layers=3, debris=deb, dim=3, seed=seed, rigid=rigid)


def dam_break_c2(seed: int = 0, rigid: bool = True) -> Scene:
    """Large single scene: the C1 tank at d = 0.004 m.

    75 x 100 x 75 = 562,500 fluid particles, three wall layers (340,908 wall
    particles, open top) and 64 debris bodies of 10x10x10 particles, one per
    cell of an 8 x 8 grid over x in [0.4, 0.8), y in [0, 0.4), with a seeded
    offset inside the cell and a seeded density in [500, 900] kg/m^3.
    """
    d = 0.004
    interior = (0.8, 0.4, 0.6)
    fluid = (0.3, 0.4, 0.3)
    off = uniform(seed, 128, stream=7, lo=0.0, hi=0.008).reshape(64, 2)
    rho = uniform(seed, 64, stream=8, lo=500.0, hi=900.0)
    deb = []
    k = 0
    for ix in range(8):
        for iy in range(8):
            x0 = 0.4 + 0.05 * ix + 0.002 + off[k, 0]
            y0 = 0.05 * iy + 0.002 + off[k, 1]
            # snap to the lattice so the body bottom layer sits d/2 above the floor
            deb.append(((round(x0 / d) * d, round(y0 / d) * d, 0.0), (10, 10, 10), float(rho[k])))
            k += 1
    return dam_break(d, interior, fluid, layers=3, debris=deb, dim=3, seed=seed, rigid=rigid)


def ensemble_c3(nscenes: int = 512, seed: int = 0, d: float = 0.01,
                interior: Sequence[float] = (0.8, 0.4, 0.6),
                fluid_range: Tuple[float, float] = (2000.0, 60000.0)) -> List[Scene]:
    """Batched ensemble: ``nscenes`` independent dam breaks in one shared tank.

    Every scene has the C1 tank (0.8 x 0.4 x 0.6 m interior, three wall
    layers, open top) at d = 0.01 m and the same periodic domain -- taken from
    the unjittered walls, so the scenes can be batched -- but its own:
      - fluid count N log-uniform in ``fluid_range`` ([2,000, 60,000]): a column spanning the tank
        depth (40 particles) whose height / width ratio is uniform in [0.5, 2];
      - 1-32 rigid cuboids with sides uniform in {3..8} particles and densities
        uniform in [400, 950] kg/m^3, in slots of a 9 d grid on the floor
        downstream of the column (at least 3 d from the fluid, beyond RadiusP),
        first layer first, then stacked layers;
      - jitter seed.
    """
    L = np.asarray(interior, float)
    nx, ny, nz = (int(round(x / d)) for x in L)
    walls = tank_walls(L, d, 3, False, 3)
    dmin, dmax = _pad_domain(walls.min(axis=0), walls.max(axis=0), d, 3.0, 3)
    out = []
    for s in range(nscenes):
        u = uniform(seed, 4, stream=1000 + 4 * s)
        lo_n, hi_n = fluid_range
        nf = float(np.exp(np.log(lo_n) + u[0] * (np.log(hi_n) - np.log(lo_n))))
        aspect = 0.5 + 1.5 * u[1]
        fy = ny
        fx = int(np.clip(round(np.sqrt(nf / (fy * aspect))), 4, nx // 2))
        fz = int(np.clip(round(nf / (fy * fx)), 4, nz - 6))
        ncub = 1 + min(int(u[2] * 32), 31)
        x0 = fx + 3
        cols = max((nx - x0 - 8) // 9 + 1, 0)   # slots whose 8-wide cuboid fits inside the tank
        rows = max((ny - 8) // 9 + 1, 0)
        slots = [(x0 + 9 * c, 9 * r, 9 * z) for z in range(3) for c in range(cols) for r in range(rows)]
        per_layer = cols * rows
        order = []
        for z in range(3):   # seeded order inside each layer, lower layers first
            perm = np.argsort(uniform(seed, per_layer, stream=1001 + 4 * s + z * 7919))
            order += [z * per_layer + int(k) for k in perm]
        ncub = min(ncub, len(order))   # small tanks hold fewer slots than cuboids
        sides = 3 + np.minimum((uniform(seed, 3 * ncub, stream=1002 + 4 * s) * 6).astype(int), 5)
        rho = uniform(seed, ncub, stream=1003 + 4 * s, lo=400.0, hi=950.0)
        deb = []
        for k in range(ncub):
            sx, sy, sz = slots[order[k]]
            deb.append(((sx * d, sy * d, sz * d), tuple(int(v) for v in sides[3 * k: 3 * k + 3]), float(rho[k])))
        sc = dam_break(d, L, (fx * d, fy * d, fz * d), layers=3, debris=deb, dim=3,
                       seed=seed * 1000003 + s, rigid=True)
        sc.domain_min, sc.domain_max = dmin.copy(), dmax.copy()
        out.append(sc)
    return out


def channel_c5(seed: int = 0, rigid: bool = True, scale: float = 1.0, nbodies: int = 2048) -> Scene:
    """Large single scene: a 1.6 x 0.8 x 0.8 m channel at d = 0.002 m.

    Three wall layers, open top (3.9M wall particles); a fluid column 0.2 x 0.8
    x 0.4 m at the x = 0 end (100 x 400 x 200 = 8.0M particles) released into
    the channel; ``nbodies`` debris cuboids with sides of U{4..12} particles per
    axis (seeded, independently) and densities U[500, 900] kg/m^3, one per slot
    of a 15 d grid over x in [0.3, 1.6) m in two layers (z = 0 and 15 d), the
    slots drawn by a seeded permutation, each body at a seeded lattice offset
    inside its slot.  ``scale`` < 1 shrinks every length (same d) and the slot
    grid with it for quick tests.
    """
    d = 0.002
    sc = scale
    interior = (1.6 * sc, 0.8 * sc, 0.8 * sc)
    fluid = (0.2 * sc, 0.8 * sc, 0.4 * sc)
    pitch = 15 * d
    x0 = 0.3 * sc
    nxs = int((interior[0] - x0) // pitch)
    nys = int(interior[1] // pitch)
    slots = [(ix, iy, iz) for iz in range(2) for ix in range(nxs) for iy in range(nys)]
    if nbodies > len(slots):
        raise ValueError("channel_c5: %d bodies, %d slots" % (nbodies, len(slots)))
    order = np.argsort(uniform(seed, len(slots), stream=9))[:nbodies]
    size = (4 + np.floor(uniform(seed, 3 * nbodies, stream=10, lo=0.0, hi=9.0 - 1e-9))).astype(int).reshape(nbodies, 3)
    off = uniform(seed, 2 * nbodies, stream=11).reshape(nbodies, 2)
    rho = uniform(seed, nbodies, stream=12, lo=500.0, hi=900.0)
    deb = []
    for b, k in enumerate(order):
        ix, iy, iz = slots[k]
        free = 15 - size[b, :2]   # lattice steps left in the slot
        ox = x0 + (ix * 15 + int(off[b, 0] * (free[0] + 1) * 0.999)) * d
        oy = (iy * 15 + int(off[b, 1] * (free[1] + 1) * 0.999)) * d
        deb.append(((ox, oy, iz * pitch), tuple(int(v) for v in size[b]), float(rho[b])))
    return dam_break(d, interior, fluid, layers=3, debris=deb, dim=3, seed=seed, rigid=rigid)


def small_2d(seed: int = 0, debris_velocity=(0.3, -0.2, 0.0), fluid_velocity=(0.0, 0.0, 0.0),
             nfx: int = 16, nfy: int = 12, d: float = 0.002, gate: bool = True) -> Scene:
    """Small 2D tank (one particle layer in z), shaped like the reference's shipped case.

    Fluid ``nfx x nfy``, a 6 x 4 debris block moving toward the fluid, a tank of
    three wall layers and, with ``gate``, a column of type-5 gate particles.
    """
    interior = ((nfx + 14) * d, (nfy + 10) * d, d)
    deb = [(((nfx + 2) * d, (nfy - 2) * d, 0.0), (6, 4, 1), 8000.0)]
    sc = dam_break(d, interior, (nfx * d, nfy * d, d), layers=3, debris=deb, dim=2, seed=seed,
                   fluid_velocity=fluid_velocity, debris_velocity=debris_velocity, max_radius_ratio=3.0)
    if gate:
        gx = (nfx + 10) * d
        gpts = lattice(np.array([gx, d * 2, 0.0]), np.array([gx + d, d * 8, d]), d, 2)
        gpts = gpts + np.c_[uniform(seed, 2 * len(gpts), 9, -0.01 * d, 0.01 * d).reshape(-1, 2), np.zeros(len(gpts))]
        sc.pos = np.concatenate([sc.pos, gpts]); sc.vel = np.concatenate([sc.vel, np.zeros_like(gpts)])
        sc.type = np.concatenate([sc.type, np.full(len(gpts), WALL_B, np.int32)])
    return sc


def params_c1(**over) -> Dict[str, object]:
    """C1 parameters: reference defaults with the workload's stated changes.

    Gravity along -z, dt = DEMDt = 1e-4 s, debris density 800 kg/m^3.  The
    debris Young's modulus is lowered from the shipped 1e11 Pa to 2e7 Pa: with
    the reference's Hertz stiffness (source/DEM-MPH.cpp:2331) a 0.02 d overlap
    gives omega*dt ~ 15 at 1e11 Pa (unstable); 2e7 Pa gives ~0.2.
    """
    p = make_params(Dt=1.0e-4, DEMDt=1.0e-4, Gravity=[0.0, 0.0, -9.81], EndTime=0.5,
                    OutputInterval=0.1, VtkOutputInterval=0.1)
    p["Density"][2] = 800.0
    p["YoungModulus"][2] = 2.0e7
    p.update(over)
    return p


def params_c3(**over) -> Dict[str, object]:
    """C3 parameters: as C1 at d = 0.01 m, dt = DEMDt = 5e-5 s (the same acoustic CFL as C1 and C2)."""
    p = params_c1(Dt=5.0e-5, DEMDt=5.0e-5, EndTime=0.15, OutputInterval=1.0, VtkOutputInterval=1.0)
    p.update(over)
    return p


def params_c5(**over) -> Dict[str, object]:
    """C5 parameters: as C1 at d = 0.002 m, dt = DEMDt = 1e-5 s (the acoustic CFL of C2), 1000 steps."""
    p = params_c1(Dt=1.0e-5, DEMDt=1.0e-5, EndTime=0.01, OutputInterval=1.0, VtkOutputInterval=1.0)
    p.update(over)
    return p


def params_c2(**over) -> Dict[str, object]:
    """C2 parameters: as C1 at dt = DEMDt = 2e-5 s (acoustic CFL c dt / d = 0.05)."""
    p = params_c1(Dt=2.0e-5, DEMDt=2.0e-5, EndTime=0.04, OutputInterval=1.0, VtkOutputInterval=1.0)
    p.update(over)
    return p
