"""Parity cases shared by the golden-fixture script and the tests.

Each case is a small seeded scene plus a parameter set chosen so that one run
exercises every branch of the step: drag in all three regimes (alpha < 0.8,
Re <= 1000, Re > 1000), buoyancy, DEM contacts solid-solid and solid-wall with
both Coulomb branches, surface tension (non-zero InteractionRatio), the
moving-wall window (0.2 <= Time <= 0.4, a rotating gate), DEM sub-steps and
neighbour rebuilds.  Debris starts at a velocity different from every fluid
neighbour's, so the drag never sees u_ij = 0 (where the reference's drag is NaN).
"""
from __future__ import annotations

import copy
import os
import sys
from dataclasses import dataclass, field
from typing import Dict, Tuple

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from dem_mph_amd import scenes as S  # noqa: E402


def _mixing_2d(seed: int) -> S.Scene:
    d = 0.002
    nfx, nfy = 14, 10
    interior = ((nfx + 14) * d, (nfy + 8) * d, d)
    # debris block overlapping the fluid's right edge zone (within RadiusP), moving
    deb = [(((nfx + 1) * d, 2 * d, 0.0), (7, 7, 1), 8000.0)]
    sc = S.dam_break(d, interior, (nfx * d, nfy * d, d), layers=3, debris=deb, dim=2, seed=seed,
                     debris_velocity=(-0.6, 0.4, 0.0), max_radius_ratio=3.0)
    nf = int(np.sum(sc.type < 2))
    u = S.uniform(seed, 2 * nf, stream=3, lo=-0.8, hi=0.8).reshape(nf, 2)
    sc.vel[:nf, :2] = u
    # a few slow fluid particles near the debris so that Re <= 1000 occurs
    sc.vel[nf - 3 * nfy: nf, :2] *= 0.01
    _embed_fluid(sc, (0.01, 0.0, 0.0), (1.0, 0.0, 0.0))
    # rotating gate (type 5)
    gx = (nfx + 9) * d
    g = S.lattice(np.array([gx, 4 * d, 0.0]), np.array([gx + 2 * d, 12 * d, d]), d, 2)
    g[:, :2] += S.uniform(seed, 2 * len(g), 9, -0.01 * d, 0.01 * d).reshape(-1, 2)
    sc.pos = np.concatenate([sc.pos, g]); sc.vel = np.concatenate([sc.vel, np.zeros_like(g)])
    sc.type = np.concatenate([sc.type, np.full(len(g), S.WALL_B, np.int32)])
    return sc


def _mixing_3d(seed: int) -> S.Scene:
    d = 0.02
    nf = (6, 5, 5)
    interior = ((nf[0] + 6) * d, 9 * d, 9 * d)
    deb = [((nf[0] * d + 0.5 * d, 1 * d, 0.0), (4, 4, 4), 800.0)]
    sc = S.dam_break(d, interior, (nf[0] * d, nf[1] * d, nf[2] * d), layers=2, debris=deb, dim=3, seed=seed,
                     debris_velocity=(-0.3, 0.2, 0.1), max_radius_ratio=3.0)
    n_f = int(np.sum(sc.type < 2))
    sc.vel[:n_f] = S.uniform(seed, 3 * n_f, stream=3, lo=-0.5, hi=0.5).reshape(n_f, 3)
    sc.vel[n_f - 25: n_f] *= 0.01
    _embed_fluid(sc, (0.01, 0.0, 0.0), (1.0, 0.0, 0.0))
    return sc


def _embed_fluid(sc, du_slow, du_fast):
    """Turn two interior debris particles into fluid particles moving at
    debris velocity + du: the surrounding solids then have alpha >= 0.8 and a
    fluid neighbour within RadiusP, with Re below and above 1000."""
    sol = np.where((sc.type >= 2) & (sc.type < 4))[0]
    c = sc.pos[sol].mean(axis=0)
    dist = np.linalg.norm(sc.pos[sol] - c, axis=1)
    picks = sol[np.argsort(dist)[[0, 3]]]
    for k, du in zip(picks, (du_slow, du_fast)):
        sc.type[k] = S.FLUID_B
        sc.vel[k] = sc.vel[k] + np.asarray(du, float)


def _params_2d(**over):
    p = S.make_params(Dt=1.0e-5, DEMDt=1.0e-5)
    p["YoungModulus"][2] = 2.0e8      # stable at dt = 1e-5 with 0.01 d jitter overlaps
    p["Wall5"] = {"Center": [0.05, 0.02, 0.001], "Velocity": [-1.0, 0.2, 0.0], "Omega": [0.0, 0.0, 30.0]}
    p.update(over)
    return p


def _params_3d(**over):
    p = S.params_c1()
    p.update(over)
    return p


def _st(p):
    r = [[0.0] * 6 for _ in range(6)]
    for a in range(6):
        for b in range(6):
            if a < 2 and b < 2:
                r[a][b] = 1.0
            elif (a < 2) != (b < 2):
                r[a][b] = 0.3 if (a < 4 and b < 4) else 0.6
    p["InteractionRatio"] = r
    return p


@dataclass
class Case:
    name: str
    dim: int
    scene: S.Scene
    params: Dict[str, object]
    steps: int
    every: int
    rigid: bool = False


def make_case(name: str) -> Case:
    if name == "mix2d":
        return Case(name, 2, _mixing_2d(11), _params_2d(), steps=30, every=15)
    if name == "st2d":
        return Case(name, 2, _mixing_2d(12), _st(_params_2d()), steps=20, every=10)
    if name == "wall2d":
        sc = _mixing_2d(13)
        sc.time = 0.2
        return Case(name, 2, sc, _params_2d(DEMDt=5.0e-6), steps=20, every=10)
    if name == "mix3d":
        return Case(name, 3, _mixing_3d(21), _params_3d(), steps=12, every=6)
    if name == "st3d":
        return Case(name, 3, _mixing_3d(22), _st(_params_3d()), steps=8, every=4)
    raise KeyError(name)


CASES = ("mix2d", "st2d", "wall2d", "mix3d", "st3d")


def write_inputs(case: Case, directory: str) -> Tuple[str, str]:
    os.makedirs(directory, exist_ok=True)
    data = os.path.join(directory, case.name + ".data")
    grid = os.path.join(directory, case.name + ".grid")
    S.write_data(case.params, data)
    S.write_grid(case.scene, grid)
    return data, grid
