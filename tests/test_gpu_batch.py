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
    if not rigid:   # loose DEM grains, as the reference moves its solid particles
        for sc in scenes:
            sc.body = None
            sc.density = None
    return scenes


@pytest.mark.parametrize("rigid", [False, True])
def test_batch_fp64_equals_each_scene_alone_bit_for_bit(rigid):
    """FP64, canonical order, rebinning every step: every field of every scene
    of the batch equals the scene's own handle bit for bit, and with rigid
    debris so does every body's state (one block per body sums its members in
    a fixed order, so the sums do not depend on the batch around them)."""
    scenes = _batch(3, 2, rigid=rigid)
    assert len({sc.n for sc in scenes}) > 1
    p = D.Params.from_dict(S.params_c1())
    kw = dict(precision=D.PREC_FP64, sort_mode=D.SORT_CANONICAL, rebuild=D.REBUILD_EVERY_STEP, rigid=rigid)
    batch = D.Solver(p, scenes, **kw)
    alone = [D.Solver(p, sc, **kw) for sc in scenes]
    info = batch.info()
    assert info.nscenes == 3 and info.n == sum(sc.n for sc in scenes)
    assert info.cells == alone[0].info().cells
    assert info.nbodies == sum(s.info().nbodies for s in alone) and (info.nbodies > 0) == rigid
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
    if rigid:
        bb = batch.bodies()
        b0 = 0
        for k, s in enumerate(alone):
            ba = s.bodies()
            nb = len(ba["mass"])
            for f in ("x", "v", "q", "l", "mass"):
                np.testing.assert_array_equal(bb[f][b0:b0 + nb], ba[f], err_msg="body %s scene %d" % (f, k))
            b0 += nb
    assert batch.time() == alone[0].time()


def test_ensemble_c3_full_size():
    """The 512-scene C3 ensemble (39.9M particles, 8,315 rigid bodies) in one
    handle, FP32 pairs with FP64 rigid bodies, rebinning every step, 5 steps.
    Size-independent properties: state finite, the block counts and every body
    mass unchanged, the clock exactly 5 dt, every scene's particles still in
    its own id range; and sampled scenes (first, last and two inside) equal,
    bit for bit, the same scene advanced in a handle of its own (a scene's rows,
    sums and body updates do not depend on the batch around it)."""
    scenes = S.ensemble_c3(512, seed=1)
    params = S.params_c3()
    kw = dict(precision=D.PREC_FP32, rigid=True, rebuild=D.REBUILD_EVERY_STEP)
    print("C3: %d scenes, %d particles" % (len(scenes), sum(sc.n for sc in scenes)), flush=True)
    batch = D.Solver(D.Params.from_dict(params), scenes, **kw)
    info = batch.info()
    assert info.nscenes == 512 and info.n == sum(sc.n for sc in scenes)
    m0 = batch.bodies()["mass"].copy()
    assert len(m0) == info.nbodies == sum(int(sc.body.max()) + 1 for sc in scenes if (sc.body >= 0).any())
    batch.step(5)
    batch.sync()
    pos, vel = batch.field("pos"), batch.field("vel")
    assert np.isfinite(pos).all() and np.isfinite(vel).all()
    t = 0.0
    for _ in range(5):   # the FP64 accumulator: 5 additions of dt from 0
        t += params["Dt"]
    assert batch.time() == t
    typ = batch.field("type", D.ORDER_SLOT)
    allt = np.concatenate([sc.type for sc in scenes])
    np.testing.assert_array_equal(np.bincount(typ, minlength=6), np.bincount(allt, minlength=6))
    np.testing.assert_array_equal(batch.bodies()["mass"], m0)
    off = batch.scene_offsets
    ids = batch.field("id", D.ORDER_SLOT)
    key = batch.field("key", D.ORDER_SLOT)
    scene_of_slot = (key % (512 * info.cells)) // info.cells
    np.testing.assert_array_equal(np.searchsorted(off, ids, side="right") - 1, scene_of_slot)
    for k in (0, 137, 300, 511):
        s = D.Solver(D.Params.from_dict(params), scenes[k], **kw)
        s.step(5)
        s.sync()
        np.testing.assert_array_equal(pos[off[k]:off[k + 1]], s.field("pos"), err_msg="pos scene %d" % k)
        np.testing.assert_array_equal(vel[off[k]:off[k + 1]], s.field("vel"), err_msg="vel scene %d" % k)
        s.close()
    batch.close()


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


def test_batch_fp32_equals_each_scene_alone_across_the_wall_motion_window():
    """FP32, rebinning every step, started 3 steps before the walls' motion
    window opens (Time 0.2 s, source/DEM-MPH.cpp:2846-2900): the walls rest
    (their slots skipped by the prediction and left out of the sort), then
    are moved by the window's transform, then rebinned with everything else.
    Every scene of the batch ends where its own handle does, field for field:
    the batch's flags and saved wall counts serve all scenes at once."""
    scenes = _batch(3, 5, rigid=True)
    prm = S.params_c1()
    for sc in scenes:
        sc.time = 0.2 - 3 * prm["Dt"]
    p = D.Params.from_dict(prm)
    kw = dict(precision=D.PREC_FP32, sort_mode=D.SORT_CANONICAL, rebuild=D.REBUILD_EVERY_STEP, rigid=True)
    batch = D.Solver(p, scenes, **kw)
    alone = [D.Solver(p, sc, **kw) for sc in scenes]
    batch.step(8)
    for s in alone:
        s.step(8)
    off = batch.scene_offsets
    for f in ("pos", "vel", "force", "vol_strain_p", "pressure_p", "kn", "neighbor_count", "id"):
        fb = batch.field(f)
        for k, s in enumerate(alone):
            ref = s.field(f)
            if f == "id":
                ref = ref + off[k]
            np.testing.assert_array_equal(fb[off[k]:off[k + 1]], ref, err_msg="%s scene %d" % (f, k))
    assert batch.time() == alone[0].time() and batch.time() > 0.2
    batch.close()
    for s in alone:
        s.close()
