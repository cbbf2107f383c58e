"""Batched scenes on the GPU (dmph_create_batch): scenes sharing a domain advance
together and each evolves as it would alone."""
import numpy as np
import pytest

import dem_mph_amd as D
from dem_mph_amd import scenes as S
import pyoracle as P

pytestmark = pytest.mark.gpu


def _batch(n, seed, rigid):
    scenes = S.ensemble_c3(n, seed=seed, d=0.02, interior=(0.6, 0.3, 0.4), fluid_range=(300.0, 3000.0))
    if not rigid:   # loose grains: the FP64 rigid-body sums use atomics, whose order is not fixed
        for sc in scenes:
            sc.body = None
            sc.density = None
    return scenes


def test_batch_fp64_equals_each_scene_alone_bit_for_bit():
    """FP64, canonical order, rebinning every step: every field of every scene
    of the batch equals the scene's own handle bit for bit."""
    scenes = _batch(3, 2, rigid=False)
    assert len({sc.n for sc in scenes}) > 1
    p = D.Params.from_dict(S.params_c1())
    kw = dict(precision=D.PREC_FP64, sort_mode=D.SORT_CANONICAL, rebuild=D.REBUILD_EVERY_STEP)
    batch = D.Solver(p, scenes, **kw)
    alone = [D.Solver(p, sc, **kw) for sc in scenes]
    info = batch.info()
    assert info.nscenes == 3 and info.n == sum(sc.n for sc in scenes)
    assert info.cells == alone[0].info().cells
    batch.step(6)
    for s in alone:
        s.step(6)
    off = batch.scene_offsets
    for f in ("pos", "vel", "force", "vol_strain_p", "divergence_p", "pressure_p", "mass", "kn", "etat", "drag_c",
              "alpha", "neighbor_count", "solid_neighbor_count", "type", "id"):
        fb = batch.field(f)
        for k, s in enumerate(alone):
            ref = s.field(f)
            if f == "id":
                ref = ref + off[k]
            np.testing.assert_array_equal(fb[off[k]:off[k + 1]], ref, err_msg="%s scene %d" % (f, k))
    assert batch.time() == alone[0].time()


def test_batch_fp32_rigid_within_stated_tolerance_of_oracle():
    """FP32 pairs with FP64 rigid bodies, one step: each scene of the batch is
    within the single-scene FP32 tolerances of the FP64 oracle run on that
    scene alone (number density 1e-5 per particle, fluid acceleration and solid
    force 1e-4 norm-relative, positions 2^-21 of a cell width, body centres 1e-9 m)."""
    scenes = _batch(3, 4, rigid=True)
    params = S.params_c1()
    batch = D.Solver(D.Params.from_dict(params), scenes, precision=D.PREC_FP32, rigid=True)
    batch.step(1)
    batch.sync()
    cw = np.array(batch.info().cell_width[:])
    off = batch.scene_offsets
    b = batch.bodies()
    boff = 0
    F, x, nd = batch.field("force"), batch.field("pos"), batch.field("vol_strain_p")
    for k, sc in enumerate(scenes):
        o = P.Oracle(P.params_from_dict(params), sc, sort_mode=P.SORT_CANONICAL, rigid=True)
        o.step(1)
        sl = slice(off[k], off[k + 1])
        n0p = o.scalars()["n0p"]
        ndr = o.field("vol_strain_p") + n0p
        assert np.max(np.abs(nd[sl] + n0p - ndr) / ndr) <= 1e-5
        typ = o.field("type")
        m = o.field("mass")
        Fr = o.field("force")
        fl, so = typ < 2, (typ >= 2) & (typ < 4)
        acc = np.linalg.norm((F[sl][fl] - Fr[fl]) / m[fl, None]) / np.linalg.norm(Fr[fl] / m[fl, None])
        assert acc <= 1e-4
        assert np.linalg.norm(F[sl][so] - Fr[so]) / np.linalg.norm(Fr[so]) <= 1e-4
        dx = x[sl] - o.field("pos")
        L = sc.domain_max - sc.domain_min
        dx -= L * np.round(dx / L)
        assert np.abs(dx).max() <= 2.0 ** -21 * cw.max()
        bx = o.get("body_x").reshape(-1, 3)
        assert np.abs(b["x"][boff:boff + len(bx)] - bx).max() <= 1e-9
        boff += len(bx)
        o.close()
    assert boff == batch.info().nbodies
