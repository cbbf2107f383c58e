"""HIP path (through the C ABI) against the reference's golden arrays and the FP64 oracle.

- FP64 + the reference's bitonic order: every slot array equals the reference's
  own (tests/golden), up to the last-bit differences of device pow/sqrt.
- FP64 / FP32 + canonical order: the oracle in the same order.
- Pair sets and cell keys bit-exact; rigid bodies against the oracle's rigid model.
Tolerances are stated per check.
"""
import os
import sys

import numpy as np
import pytest

import cases as CS
import golden_io as G
import pyoracle as P
import dem_mph_amd as D
from dem_mph_amd import scenes as S

pytestmark = pytest.mark.gpu

# golden name -> (library field, components)
GOLDEN_FIELDS = {
    "ParticleIndex": "id", "Property": "type", "CellIndex": "key", "NeighborCount": "neighbor_count",
    "SolidNeighborCount": "solid_neighbor_count", "Position": "pos", "Velocity": "vel", "Force": "force",
    "DensityA": "density_a", "GravityCenter": "gravity_center", "PressureA": "pressure_a",
    "VolStrainP": "vol_strain_p", "DivergenceP": "divergence_p", "PressureP": "pressure_p", "Kappa": "kappa",
    "Mass": "mass", "alpha": "alpha", "DragC": "drag_c", "NormalSpringConstant": "kn",
    "NormalViscousDampingConstant": "etan", "TangentialSpringConstant": "kt",
    "TangentialViscousDampingConstant": "etat", "VirialPressureAtParticle": "virial_pressure",
}


def _solver(g, **kw):
    return D.Solver(D.Params.from_dict(G.params(g)), G.scene(g), **kw)


def _relerr(a, b, floor):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)) if a.size else 0.0


def _report(name, errs):
    sys.stdout.write("\n[%s] " % name + " ".join("%s=%.2e" % kv for kv in sorted(errs.items())))
    sys.stdout.flush()


@pytest.mark.parametrize("name", CS.CASES)
def test_fp64_reference_order_matches_golden(name):
    """FP64, the reference's sort: slot arrays vs the reference's after N steps.
    Integers exact; reals within 1e-9 relative (scale floor 1e-6 of the field's
    largest magnitude): only device pow/log last bits may differ."""
    g = G.load(name)
    print("case", name, flush=True)
    s = _solver(g, precision=D.PREC_FP64, sort_mode=D.SORT_REFERENCE)
    # step 0: permutation, keys and the CSR neighbour list bit for bit
    np.testing.assert_array_equal(s.field("id", D.ORDER_SLOT), g["s0_ParticleIndex"])
    np.testing.assert_array_equal(s.field("key", D.ORDER_SLOT), g["s0_CellIndex"])
    ptr, idx = s.pairs()
    np.testing.assert_array_equal(ptr, g["s0_NeighborPtr"])
    np.testing.assert_array_equal(idx, g["s0_NeighborInd"])
    np.testing.assert_array_equal(s.field("vol_strain_p", D.ORDER_SLOT), g["s0_VolStrainP"])
    s.step(G.steps(g))
    s.virial()
    s.sync()
    typ = g["sN_Property"]
    errs = {}
    for gf, lf in GOLDEN_FIELDS.items():
        ref = g["sN_" + gf]
        got = s.field(lf, D.ORDER_SLOT)
        if ref.dtype.kind == "i":
            np.testing.assert_array_equal(got.reshape(ref.shape), ref, err_msg=gf)
            continue
        got = got.reshape(-1)
        if gf == "Force":   # wall forces are dead in the reference and not computed here
            keep = np.repeat(typ < 4, 3)
            got, ref = got[keep], ref[keep]
        floor = 1e-6 * max(np.abs(ref).max(), 1e-300)
        errs[gf] = _relerr(got, ref, floor)
    _report(name, errs)
    bad = {k: v for k, v in errs.items() if not v <= 1e-9}
    assert not bad, bad
    assert abs(s.time() - float(g["scalars"][0])) == 0.0


@pytest.mark.parametrize("name", ["mix2d", "st3d"])
def test_fp64_canonical_order_matches_oracle(name):
    g = G.load(name)
    s = _solver(g, precision=D.PREC_FP64, sort_mode=D.SORT_CANONICAL)
    o = P.Oracle(P.params_from_dict(G.params(g)), G.scene(g), sort_mode=P.SORT_CANONICAL)
    np.testing.assert_array_equal(s.field("id", D.ORDER_SLOT), o.get("id"))
    ptr, idx = s.pairs()
    np.testing.assert_array_equal(ptr, o.get("neighbor_ptr"))
    np.testing.assert_array_equal(idx, o.get("neighbor_ind"))
    k = G.steps(g)
    s.step(k); o.step(k)
    s.virial(); o.virial()
    np.testing.assert_array_equal(s.field("id", D.ORDER_SLOT), o.get("id"))
    errs = {}
    for lf in ("pos", "vel", "density_a", "vol_strain_p", "divergence_p", "pressure_p", "drag_c", "kn", "kt",
               "virial_pressure", "alpha"):
        ref = o.field(lf, original_order=False).reshape(-1)
        floor = 1e-6 * max(np.abs(ref).max(), 1e-300)
        errs[lf] = _relerr(s.field(lf, D.ORDER_SLOT).reshape(-1), ref, floor)
    _report(name, errs)
    assert max(errs.values()) <= 1e-9, errs


def _case(name):
    """(scene, parameter dict, domain length) of a golden case or of C1 at full size."""
    if name == "c1":
        sc = S.dam_break_c1(seed=0)
        return sc, S.params_c1(), sc.domain_max - sc.domain_min
    g = G.load(name)
    return G.scene(g), G.params(g), g["domain_max"] - g["domain_min"]


def _fp32_vs_oracle(name, steps, compensated=True, lanes=0):
    sc, prm, L = _case(name)
    s = D.Solver(D.Params.from_dict(prm), sc, precision=D.PREC_FP32, sort_mode=D.SORT_CANONICAL,
                 compensated=compensated, lanes=lanes)
    o = P.Oracle(P.params_from_dict(prm), sc, sort_mode=P.SORT_CANONICAL)
    # binning happens in FP64 on the first call: identical keys and permutation
    np.testing.assert_array_equal(s.field("key", D.ORDER_SLOT), o.get("key"))
    np.testing.assert_array_equal(s.field("id", D.ORDER_SLOT), o.get("id"))
    s.step(steps); o.step(steps)
    return L, s, o


@pytest.mark.parametrize("name,compensated,lanes", [(n, c, 8) for n in ("c1", "mix3d", "st3d", "mix2d", "st2d", "wall2d")
                                                   for c in (True, False)] +
                         [(n, True, l) for n in ("mix3d", "st3d", "mix2d") for l in (2, 4, 16)])
def test_fp32_one_step_within_stated_tolerance(name, compensated, lanes):
    """FP32 pairs on cell-relative positions, one step, with FP64 or FP32 sums
    and 2, 4, 8 or 16 lanes per particle, against the FP64 oracle.
    Stated tolerances:
      number density sum_j w_p (= VolStrainP + N0p), per particle     1e-5 relative
      fluid acceleration F/m: over the field (||da|| / ||a||)         1e-5
                              per particle (|da_i| / |a_i|)           2e-4
      solid (contact + drag + buoyancy + weight) force:
                              over the field                          5e-5
                              per particle, |dF_i| / max(|F_i|, median_j |F_j|)   1e-3
      positions                                   2^-21 x CellWidth (8 ulp of an FP32 offset)
    The per-particle acceleration and contact-force bounds are looser than
    1e-5 because positions are stored as FP32 offsets: the oracle run on the
    GPU's rounded initial positions differs from the oracle on the exact ones
    by up to 4e-5 per particle in acceleration (C1) and 3e-4 in solid force --
    the overlap d - r of a contact and the small VolStrainP = sum w - N0p of a
    fluid particle amplify a 4e-9 m rounding of a position."""
    L, s, o = _fp32_vs_oracle(name, 1, compensated, lanes)
    assert tuple(s.info().lanes) == (lanes, lanes)
    n0p = o.scalars()["n0p"]
    nd_ref = o.field("vol_strain_p") + n0p
    nd = s.field("vol_strain_p") + n0p
    per = np.max(np.abs(nd - nd_ref) / np.abs(nd_ref))
    typ = o.field("type")
    F, Fr = s.field("force"), o.field("force")
    m = o.field("mass")
    fl = typ < 2
    a, ar = F[fl] / m[fl, None], Fr[fl] / m[fl, None]
    acc_err = np.linalg.norm(a - ar) / np.linalg.norm(ar)
    acc_pp = np.max(np.linalg.norm(a - ar, axis=1) / np.linalg.norm(ar, axis=1))
    sol = (typ >= 2) & (typ < 4)
    sf_err = sf_pp = 0.0
    if sol.any():
        nr = np.linalg.norm(Fr[sol], axis=1)
        sf_err = np.linalg.norm(F[sol] - Fr[sol]) / np.linalg.norm(Fr[sol])
        sf_pp = np.max(np.linalg.norm(F[sol] - Fr[sol], axis=1) / np.maximum(nr, np.median(nr)))
    dx = s.field("pos") - o.field("pos")
    dx = dx - L * np.round(dx / L)
    _report("%s-%s-%d" % (name, "fp64sums" if compensated else "fp32sums", lanes),
            {"density_per_particle": per, "fluid_accel": acc_err, "fluid_accel_per_particle": acc_pp,
             "solid_force": sf_err, "solid_force_per_particle": sf_pp, "pos_abs": np.abs(dx).max()})
    assert per <= 1e-5
    assert acc_err <= 1e-5 and acc_pp <= 2e-4
    assert sf_err <= 5e-5 and sf_pp <= 1e-3
    cw = np.array(s.info().cell_width[:])
    assert np.abs(dx).max() <= 2.0 ** -21 * cw.max()


@pytest.mark.parametrize("rigid", [False, True])
def test_c1_dam_break_fp64_matches_oracle(rigid):
    """C1 at its own size (4,500 fluid, 256 debris, 15,468 wall particles,
    d = 0.02 m) in FP64, canonical order, the reference's 0.05 d skin schedule,
    300 steps (0.03 s) against the FP64 oracle: the slot order equal and every
    field within 1e-9 relative (scale floor 1e-6 of the field's largest value;
    wall forces are not maintained).  Loose grains are the reference's model;
    rigid bodies add the body state (X within 1e-12 m, q within 1e-12)."""
    sc = S.dam_break_c1(seed=0, rigid=rigid)
    prm = S.params_c1()
    s = D.Solver(D.Params.from_dict(prm), sc, precision=D.PREC_FP64, rigid=rigid)
    o = P.Oracle(P.params_from_dict(prm), sc, sort_mode=P.SORT_CANONICAL, rigid=rigid)
    print("C1 rigid=%d: %d particles" % (rigid, sc.n), flush=True)
    errs = {}
    for k in (100, 200):
        s.step(k); o.step(k)
        s.sync()
        np.testing.assert_array_equal(s.field("id", D.ORDER_SLOT), o.get("id"))
        typ = o.field("type")
        for f in ("pos", "vel", "force", "vol_strain_p", "pressure_p", "divergence_p", "kn", "drag_c"):
            a, b = s.field(f), o.field(f)
            if f == "force":
                a, b = a[typ < 4], b[typ < 4]
            errs[f] = max(errs.get(f, 0.0), _relerr(a, b, 1e-6 * max(np.abs(b).max(), 1e-300)))
    if rigid:
        b = s.bodies()
        errs["body_x_abs"] = np.abs(b["x"] - o.get("body_x").reshape(-1, 3)).max()
        errs["body_q_abs"] = np.abs(b["q"] - o.get("body_q").reshape(-1, 4)).max()
        assert errs.pop("body_x_abs") <= 1e-12 and errs.pop("body_q_abs") <= 1e-12
    _report("c1-fp64-rigid%d" % rigid, errs)
    assert s.info().rebuilds > 1
    assert max(errs.values()) <= 1e-9, errs
    assert s.time() == o.scalars()["time"]
    s.close(); o.close()


@pytest.mark.parametrize("rigid", [False, True])
def test_c1_dam_break_fp32_stays_close(rigid):
    """C1 at full size in FP32 (mixed precision), 300 steps (0.03 s) against
    the FP64 oracle, the same 0.05 d skin schedule (the rebuild count must
    match).  Stated bounds: fluid positions within 1e-6 m (5e-5 d), fluid
    velocities within 1e-4 norm-relative (measured 6.4e-7 m, 5e-6); rigid
    debris positions within 1e-6 m (measured 3.2e-7 m); total mass of the
    moving particles within 1e-6 relative.  Loose DEM grains (the reference's
    own debris model) are not bounded: with their stiff Hertz contacts and the
    coefficients of the last contact in row order, rounding differences grow
    to ~2 mm over the 300 steps (reported)."""
    sc = S.dam_break_c1(seed=0, rigid=rigid)
    prm = S.params_c1()
    L = sc.domain_max - sc.domain_min
    s = D.Solver(D.Params.from_dict(prm), sc, precision=D.PREC_FP32, rigid=rigid)
    o = P.Oracle(P.params_from_dict(prm), sc, sort_mode=P.SORT_CANONICAL, rigid=rigid)
    s.step(300); o.step(300)
    s.sync()
    assert s.info().rebuilds == int(o.scalars()["rebuilds"])
    dx = s.field("pos") - o.field("pos")
    dx -= L * np.round(dx / L)
    v, vr = s.field("vel"), o.field("vel")
    typ = o.field("type")
    fl, so, mv = typ < 2, (typ >= 2) & (typ < 4), typ < 4
    m, mr = s.field("mass"), o.field("mass")
    errs = {"fluid_pos_abs": np.abs(dx[fl]).max(), "fluid_vel_rel": np.linalg.norm(v[fl] - vr[fl]) / np.linalg.norm(vr[fl]),
            "debris_pos_abs": np.abs(dx[so]).max(), "mass_rel": abs(m[mv].sum() - mr[mv].sum()) / mr[mv].sum()}
    _report("c1-fp32-300-rigid%d" % rigid, errs)
    assert errs["fluid_pos_abs"] <= 1e-6 and errs["fluid_vel_rel"] <= 1e-4 and errs["mass_rel"] <= 1e-6
    assert np.isfinite(dx).all()
    if rigid:
        assert errs["debris_pos_abs"] <= 1e-6
    s.close(); o.close()


@pytest.mark.parametrize("name", ["mix3d"])
def test_fp32_multi_step_stays_close(name):
    L, s, o = _fp32_vs_oracle(name, G.steps(G.load(name)))
    dx = s.field("pos") - o.field("pos")
    dx = dx - L * np.round(dx / L)
    v, vr = s.field("vel"), o.field("vel")
    verr = np.linalg.norm(v - vr) / np.linalg.norm(vr)
    _report(name + "-multi", {"pos_abs": np.abs(dx).max(), "vel_rel": verr})
    assert np.abs(dx).max() <= 1e-7
    assert verr <= 1e-3


@pytest.mark.parametrize("start", [0.2, 0.4])
def test_fp32_walls_through_the_motion_window_edge(start):
    """wall2d's walls move while 0.2 <= Time <= 0.4 s (source/DEM-MPH.cpp:
    2846-2900).  Started 3 steps before an edge of that window and run 8 steps
    with the FP64 oracle beside it: the walls start moving, or stop and are
    zeroed, then stand still (the FP32 step skips walls at rest in the
    prediction, and leaves static walls out of the sort).  Every wall position
    matches the oracle's to FP32 rounding of the cell-relative offsets, and
    the state stays as close as over the golden run."""
    sc, prm, L = _case("wall2d")
    sc.time = start - 3 * prm["Dt"]
    s = D.Solver(D.Params.from_dict(prm), sc, precision=D.PREC_FP32, sort_mode=D.SORT_CANONICAL)
    o = P.Oracle(P.params_from_dict(prm), sc, sort_mode=P.SORT_CANONICAL)
    s.step(8); o.step(8)
    typ = o.field("type")
    wall = typ >= 4
    assert wall.any()
    dx = s.field("pos") - o.field("pos")
    dx = dx - L * np.round(dx / L)
    v, vr = s.field("vel"), o.field("vel")
    verr = np.linalg.norm(v - vr) / np.linalg.norm(vr)
    cw = np.array(s.info().cell_width[:])
    _report("wall2d-edge-%g" % start, {"wall_pos_abs": np.abs(dx[wall]).max(), "pos_abs": np.abs(dx).max(),
                                       "vel_rel": verr})
    assert np.abs(dx[wall]).max() <= 2.0 ** -21 * cw.max()
    assert np.abs(dx).max() <= 1e-7
    assert verr <= 1e-3
    np.testing.assert_array_equal(v[wall] == 0.0, vr[wall] == 0.0)
    s.close(); o.close()


def _rigid_scene():
    g = G.load("mix3d")
    sc = G.scene(g)
    sol = (sc.type >= 2) & (sc.type < 4)
    sc.body = np.where(sol, 0, -1).astype(np.int32)
    sc.density = np.where(sol, 750.0, -1.0)
    return g, sc


@pytest.mark.parametrize("prec,body,tol,tolx", [(D.PREC_FP64, D.PREC_FP64, 1e-10, 1e-12),
                                                (D.PREC_FP32, D.PREC_FP64, 2e-5, 1e-9),
                                                (D.PREC_FP32, D.PREC_FP32, 1e-4, 1e-7)])
def test_rigid_bodies_match_oracle(prec, body, tol, tolx):
    """Body state after 12 steps against the oracle's FP64 rigid model.
    FP64: 1e-10 relative on V, L, q and 1e-12 m on X.  Mixed (FP32 pairs,
    FP64 body sums and integration): 2e-5 relative, X 1e-9 m.  Plain FP32
    (body sums and integration in FP32 too): 1e-4 relative, X 1e-7 m."""
    g, sc = _rigid_scene()
    p = G.params(g)
    s = D.Solver(D.Params.from_dict(p), sc, precision=prec, rigid=True, rigid_precision=body)
    o = P.Oracle(P.params_from_dict(p), sc, sort_mode=P.SORT_CANONICAL, rigid=True)
    s.step(12); o.step(12)
    b = s.bodies()
    np.testing.assert_allclose(b["mass"], o.get("body_mass"), rtol=1e-14)
    X = o.get("body_x").reshape(-1, 3); V = o.get("body_v").reshape(-1, 3); Lm = o.get("body_l").reshape(-1, 3)
    q = o.get("body_q").reshape(-1, 4)
    errs = {"X": np.abs(b["x"] - X).max(), "V": np.abs(b["v"] - V).max() / np.abs(V).max(),
            "L": np.abs(b["l"] - Lm).max() / max(np.abs(Lm).max(), 1e-30), "q": np.abs(b["q"] - q).max()}
    _report("rigid-%d-body%d" % (prec, body), errs)
    assert errs["X"] <= tolx
    assert errs["V"] <= tol and errs["L"] <= max(tol, 1e-8) and errs["q"] <= max(tol, 1e-9)


@pytest.mark.parametrize("name,block,cell_scale", [("mix3d", 64, 1.0), ("st3d", 128, 1.0), ("mix2d", 64, 1.0),
                                                   ("mix3d", 256, 1.5), ("c1", 128, 1.5), ("st2d", 256, 1.25)])
def test_fp32_workgroups_and_cell_sizes(name, block, cell_scale):
    """The autotuner's other parameters: sweeps in workgroups of 64 or 128
    threads, and cells wider than MaxRadius + 0.1 d (cell_scale).  Rows still
    hold exactly the FP64 rows' pairs at the first binning (the cutoff test is
    exact and the 3^D stencil of a wider cell still covers it), and one step
    stays within the FP32 tolerances against the FP64 oracle on the reference's
    own cell grid (number density 1e-5 per particle, fluid acceleration 1e-5
    over the field, solid force 5e-5 over the field).  Wider cells hold
    larger FP32 offsets, whose rounding grows with them: fluid acceleration
    within 5e-5 there (measured 1.7e-5 on C1 at 1.5x).  And the solid force is
    not held to a bound: a DEM particle takes the coefficients of its last
    contact in row order (source/DEM-MPH.cpp:2331-2396) and rows follow the
    cell grid, so another contact's coefficients can win; the force stays
    finite and is reported."""
    sc, prm, L = _case(name)
    s = D.Solver(D.Params.from_dict(prm), sc, precision=D.PREC_FP32, block=block, cell_scale=cell_scale)
    info = s.info()
    assert tuple(info.block) == (block, block) and abs(info.cell_scale - cell_scale) < 1e-12
    s64 = D.Solver(D.Params.from_dict(prm), sc, precision=D.PREC_FP64)
    cw, cw64 = np.array(info.cell_width[:]), np.array(s64.info().cell_width[:])
    # cells at least cell_scale x (MaxRadius + 0.1 d) wide: never narrower than the reference's
    assert np.all(cw[:sc.dim] >= cell_scale * (info.max_radius + 0.1 * sc.d) * (1 - 1e-12))
    assert np.all(cw[:sc.dim] >= cw64[:sc.dim] * (1 - 1e-12))
    p32, i32 = s.pairs()
    p64, i64 = s64.pairs()
    id32, id64 = s.field("id", D.ORDER_SLOT), s64.field("id", D.ORDER_SLOT)
    # rows as sets of original ids, keyed by original id
    r32 = {int(id32[k]): frozenset(id32[i32[p32[k]:p32[k + 1]]].tolist()) for k in range(len(id32))}
    r64 = {int(id64[k]): frozenset(id64[i64[p64[k]:p64[k + 1]]].tolist()) for k in range(len(id64))}
    assert r32 == r64
    s64.close()
    o = P.Oracle(P.params_from_dict(prm), sc, sort_mode=P.SORT_CANONICAL)
    s.step(1); o.step(1)
    n0p = o.scalars()["n0p"]
    nd, ndr = s.field("vol_strain_p") + n0p, o.field("vol_strain_p") + n0p
    typ = o.field("type")
    F, Fr, m = s.field("force"), o.field("force"), o.field("mass")
    fl, so = typ < 2, (typ >= 2) & (typ < 4)
    errs = {"density_per_particle": np.max(np.abs(nd - ndr) / np.abs(ndr)),
            "fluid_accel": np.linalg.norm((F[fl] - Fr[fl]) / m[fl, None]) / np.linalg.norm(Fr[fl] / m[fl, None]),
            "solid_force": np.linalg.norm(F[so] - Fr[so]) / np.linalg.norm(Fr[so]) if so.any() else 0.0}
    _report("%s-block%d-cells%.2f" % (name, block, cell_scale), errs)
    assert errs["density_per_particle"] <= 1e-5 and errs["fluid_accel"] <= (1e-5 if cell_scale == 1.0 else 5e-5)
    assert np.isfinite(F).all()
    if cell_scale == 1.0:
        assert errs["solid_force"] <= 5e-5
    s.close(); o.close()


@pytest.mark.parametrize("wall_lanes", [1, 2, 4])
def test_static_wall_density_between_full_sweeps(wall_lanes):
    """Static walls keep their wall neighbours' density sum from the last full
    sweep and walk only their rows' fluid and solid entries on the other steps.
    mix3d's walls stand still; the first two steps sweep everything, steps 3-6
    do not (no rebinning: the skin holds).  After every step the walls' number
    density stays within 1e-5 per particle of the FP64 oracle, as the fluid's,
    and the walls' PressureP within 1e-4 norm-relative -- with 1, 2 or 4 lanes
    per wall row in pass A (the fluid rows keep their own lane count)."""
    g = G.load("mix3d")
    sc, prm = G.scene(g), G.params(g)
    s = D.Solver(D.Params.from_dict(prm), sc, precision=D.PREC_FP32, sort_mode=D.SORT_CANONICAL, rigid=False,
                 wall_lanes=wall_lanes)
    assert s.info().wall_lanes == wall_lanes
    o = P.Oracle(P.params_from_dict(prm), sc, sort_mode=P.SORT_CANONICAL)
    n0p = o.scalars()["n0p"]
    typ = o.field("type")
    wall = typ >= 4
    assert wall.any()
    reb2 = None
    for k in range(6):
        s.step(1); o.step(1)
        if k == 1:
            reb2 = s.info().rebuilds
        nd, ndr = s.field("vol_strain_p") + n0p, o.field("vol_strain_p") + n0p
        rel = np.abs(nd - ndr) / np.abs(ndr)
        pw, pwr = s.field("pressure_p")[wall], o.field("pressure_p")[wall]
        prel = np.linalg.norm(pw - pwr) / max(np.linalg.norm(pwr), 1e-300)
        assert rel[wall].max() <= 1e-5 and rel.max() <= 1e-5, (k, rel[wall].max(), rel.max())
        assert prel <= 1e-4, (k, prel)
    assert s.info().rebuilds == reb2   # steps 3-6 rebinned nothing: the partial wall walk ran
    s.close(); o.close()


def test_static_wall_dem_coefficients_kept():
    """A static wall whose last DEM contact is a wall keeps its coefficients
    between full sweeps instead of recomputing them.  Two identical FP32 runs
    of mix3d; after 6 steps the second writes its positions back (a rebinning
    and full sweeps follow), the first keeps going: after one more step every
    wall's four coefficients are equal bit for bit."""
    g = G.load("mix3d")
    sc, prm = G.scene(g), G.params(g)
    a = D.Solver(D.Params.from_dict(prm), sc, precision=D.PREC_FP32)
    b = D.Solver(D.Params.from_dict(prm), sc, precision=D.PREC_FP32)
    a.step(6); b.step(6)
    a.sync(); b.sync()
    b.set_field("pos", b.field("pos"))
    a.step(1); b.step(1)
    a.sync(); b.sync()
    assert b.info().rebuilds == a.info().rebuilds + 1
    wall = a.field("type") >= 4
    for f in ("kn", "etan", "kt", "etat"):
        np.testing.assert_array_equal(a.field(f)[wall], b.field(f)[wall], err_msg=f)
    a.close(); b.close()


@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_pass_b_staged_in_lds_equals_global_gathers(monkeypatch, lanes):
    """Pass B with its neighbour cells staged in LDS (DMPH_PASS_B=lds: a
    workgroup copies the 27 slot runs around a tile of cells, then gathers from
    LDS) walks the same entries in the same order with the same arithmetic as
    the default pass B: after 40 steps of C1 at full size (rebuilds included)
    and of mix3d, every field equals the default run's bit for bit (tiles at
    the grid's ends, and tiles with more records than the buffer, take the
    kernel's global-gather fallback)."""
    for sc, prm in ((S.dam_break_c1(seed=3, rigid=True), S.params_c1()),
                    (G.scene(G.load("mix3d")), G.params(G.load("mix3d")))):
        out = []
        for mode in ("", "lds"):
            monkeypatch.setenv("DMPH_PASS_B", mode)
            s = D.Solver(D.Params.from_dict(prm), sc, precision=D.PREC_FP32, rigid=sc.body is not None, lanes=lanes)
            s.step(40)
            s.sync()
            out.append({f: s.field(f) for f in ("pos", "vel", "force", "pressure_p", "vol_strain_p")})
            s.close()
        for f in out[0]:
            np.testing.assert_array_equal(out[1][f], out[0][f], err_msg=f)


def test_cell_size_tuned_at_creation():
    """cell_scale < 0 tunes the cell width at creation (FP32): the handle comes
    back on one of the candidate grids (1, 1.25 or 1.5 x the reference's cell),
    at its initial state -- the candidates' trial steps leave no trace: the
    same step count and state as a handle created directly on that grid -- and
    FP64 keeps the reference's grid."""
    sc, prm, L = _case("mix3d")
    p = D.Params.from_dict(prm)
    s = D.Solver(p, sc, precision=D.PREC_FP32, cell_scale=-1.0)
    c = float(s.info().cell_scale)
    assert min(abs(c - x) for x in (1.0, 1.25, 1.5)) < 1e-12
    ref = D.Solver(p, sc, precision=D.PREC_FP32, cell_scale=c)
    assert s.time() == ref.time() == sc.time
    for k in (1, 3):
        s.step(k); ref.step(k)
        s.sync(); ref.sync()
        for f in ("pos", "vel", "force", "key"):
            np.testing.assert_array_equal(s.field(f), ref.field(f), err_msg=f)
    s.close(); ref.close()
    s64 = D.Solver(p, sc, precision=D.PREC_FP64, cell_scale=-1.0)
    assert float(s64.info().cell_scale) == 1.0
    s64.close()


@pytest.mark.parametrize("prec", [D.PREC_FP64, D.PREC_FP32])
def test_rigid_body_pose_over_1000_steps(prec):
    """The debris of mix3d as one rigid body in the fluid (d = 0.02 m), 1000
    steps (0.1 s) against the oracle's rigid model.  Stated tolerances on the
    body's centre X and orientation quaternion q (absolute, max over components
    and over steps 250, 500, 1000):
      FP64 (device pow/log/sin/cos last bits only)      X 1e-12 m,          q 1e-12
      FP32 pairs, FP64 body reductions and integration  X 1e-4 m (5e-3 d),  q 1e-3
    Measured: FP64 1.1e-15 m and 1.0e-14.  FP32: between 7e-9 m / 2e-7 and
    3e-6 m / 5e-5 depending on the FP32 sweeps' rounding -- the body keeps
    hitting walls and grains, and the DEM contact model (coefficients of the
    last contact in row order, the Coulomb cap's branch) switches at
    thresholds, so rounding differences grow by orders of magnitude at some
    contacts."""
    g, sc = _rigid_scene()
    p = G.params(g)
    s = D.Solver(D.Params.from_dict(p), sc, precision=prec, rigid=True)
    o = P.Oracle(P.params_from_dict(p), sc, sort_mode=P.SORT_CANONICAL, rigid=True)
    tx, tq = (1e-12, 1e-12) if prec == D.PREC_FP64 else (1e-4, 1e-3)
    errs = {}
    done = 0
    for k in (250, 250, 500):
        s.step(k); o.step(k)
        done += k
        b = s.bodies()
        ex = np.abs(b["x"] - o.get("body_x").reshape(-1, 3)).max()
        eq = np.abs(b["q"] - o.get("body_q").reshape(-1, 4)).max()
        errs["X_%d" % done] = ex; errs["q_%d" % done] = eq
        errs["X"] = max(errs.get("X", 0.0), ex)
        errs["q"] = max(errs.get("q", 0.0), eq)
    _report("rigid-1000-%d" % prec, errs)
    assert s.time() == o.scalars()["time"]
    assert errs["X"] <= tx and errs["q"] <= tq
    # the body has moved and turned by far more than the tolerance
    q = s.bodies()["q"][0]
    assert abs(q[0]) < 1.0 - 1e-6
    s.close(); o.close()


def test_rigid_body_free_fall_on_gpu():
    d = 0.02
    pts = S.lattice(np.full(3, 10 * d), np.full(3, 13 * d), d, 3)
    sc = S.Scene(d=d, dim=3, domain_min=np.zeros(3), domain_max=np.full(3, 40 * d), type=np.full(len(pts), 2, np.int32),
                 pos=pts, vel=np.zeros_like(pts), body=np.zeros(len(pts), np.int32), density=np.full(len(pts), 700.0))
    p = S.params_c1()
    s = D.Solver(D.Params.from_dict(p), sc, precision=D.PREC_FP64, rigid=True)
    x0 = s.bodies()["x"][0, 2]
    n = 20
    s.step(n)
    dt = p["Dt"]
    assert abs(s.bodies()["x"][0, 2] - (x0 - 9.81 * dt * dt * n * (n + 1) / 2)) < 1e-12


def test_large_scene_properties_fp32():
    """The C2 tank at d = 0.004 m (about 0.97M particles), FP32, rigid debris:
    cell keys after the first binning equal the FP64 formula; after 5 steps
    there are no NaNs, particle counts per block are unchanged, body masses are
    conserved and the fluid has started to fall."""
    sc = S.dam_break_c2(seed=1)
    p = S.params_c2()
    s = D.Solver(D.Params.from_dict(p), sc, precision=D.PREC_FP32, rigid=True)
    info = s.info()
    assert info.n == sc.n and info.nbodies == 64
    cc = np.array(info.cell_count[:]); cw = np.array(info.cell_width[:])
    c = np.floor((sc.pos - sc.domain_min) / cw).astype(np.int64) % cc
    key = (c[:, 0] * cc[1] + c[:, 1]) * cc[2] + c[:, 2]
    blk = np.where(sc.type < 2, 0, np.where(sc.type < 4, 1, 2))
    np.testing.assert_array_equal(s.field("key"), key + blk * info.cells)
    m0 = s.bodies()["mass"].copy()
    s.step(5)
    s.sync()
    v = s.field("vel")
    assert np.isfinite(v).all() and np.isfinite(s.field("pos")).all()
    np.testing.assert_array_equal(np.sort(s.field("type", D.ORDER_SLOT)), np.sort(sc.type))
    np.testing.assert_allclose(s.bodies()["mass"], m0, rtol=0)
    fl = sc.type < 2
    assert v[fl, 2].mean() < 0.0


@pytest.mark.parametrize("prec,rebuild", [(D.PREC_FP32, D.REBUILD_EVERY_STEP), (D.PREC_FP32, D.REBUILD_SKIN),
                                          (D.PREC_FP64, D.REBUILD_EVERY_STEP)])
def test_graph_replays_match_launches(prec, rebuild):
    """The step captured as a hipGraph and replayed, with host reads and
    synchronisations between replays, leaves the same state bit for bit as the
    same steps launched one by one (with rebinning every step, every replay
    runs the sort and the list build)."""
    g = G.load("mix3d")
    eager = _solver(g, precision=prec, rebuild=rebuild, graph=False)
    graph = _solver(g, precision=prec, rebuild=rebuild, graph=True)
    fields = ("pos", "vel", "force", "pressure_p", "vol_strain_p", "kn", "drag_c", "key")
    for k in (1, 1, 1, 3, 7):
        eager.step(k)
        graph.step(k)
        eager.sync()
        graph.sync()
        for f in fields:
            np.testing.assert_array_equal(graph.field(f), eager.field(f), err_msg=f)
    assert graph.info().rebuilds == eager.info().rebuilds
    if rebuild == D.REBUILD_EVERY_STEP:
        assert graph.info().rebuilds > 13
    # switching to per-group timing (launches one by one) and back keeps the state in step
    graph.set_timing(True); graph.step(2); graph.set_timing(False); graph.step(2)
    eager.step(4)
    graph.sync(); eager.sync()
    np.testing.assert_array_equal(graph.field("pos"), eager.field("pos"))


@pytest.mark.parametrize("prec", [D.PREC_FP64, D.PREC_FP32])
def test_rigid_bodies_are_reproducible(prec):
    """Rigid bodies, rebinning every step: graph replays and one-by-one launches
    leave the same particle and body state bit for bit (the per-body sums run
    in a fixed order, no atomics), and so does a second identical run."""
    g, sc = _rigid_scene()
    p = D.Params.from_dict(G.params(g))
    runs = [D.Solver(p, sc, precision=prec, rigid=True, rebuild=D.REBUILD_EVERY_STEP, graph=gr)
            for gr in (False, True, True)]
    for k in (1, 2, 5):
        for s in runs:
            s.step(k)
            s.sync()
        for s in runs[1:]:
            for f in ("pos", "vel", "force"):
                np.testing.assert_array_equal(s.field(f), runs[0].field(f), err_msg=f)
            b, b0 = s.bodies(), runs[0].bodies()
            for f in ("x", "v", "q", "l"):
                np.testing.assert_array_equal(b[f], b0[f], err_msg="body " + f)


@pytest.mark.parametrize("prec", [D.PREC_FP64, D.PREC_FP32])
def test_set_field_positions_and_velocities(prec):
    """dmph_set_field: values read back as written (positions wrapped into the
    periodic domain), the next step rebins unconditionally and the fluid's
    cell keys follow the new positions."""
    g = G.load("mix3d")
    s = _solver(g, precision=prec)
    s.step(2)
    s.sync()
    info = s.info()
    cw, cc = np.array(info.cell_width[:]), np.array(info.cell_count[:])
    lo, hi = np.asarray(g["domain_min"], float), np.asarray(g["domain_max"], float)
    L = hi - lo
    typ = s.field("type")
    fl = typ < 2
    x2 = s.field("pos")
    x2[fl, 0] += 1.5 * cw[0]          # every fluid particle moves one and a half cells
    v2 = 2.0 * s.field("vel")
    s.set_field("pos", x2)
    s.set_field("vel", v2)
    dx = s.field("pos") - x2
    dx -= L * np.round(dx / L)
    tol = 1e-12 if prec == D.PREC_FP64 else 2.0 ** -21 * cw.max()
    assert np.abs(dx).max() <= tol
    np.testing.assert_allclose(s.field("vel"), v2, rtol=0 if prec == D.PREC_FP64 else 1e-7, atol=0)
    r0 = s.info().rebuilds
    s.step(1)
    s.sync()
    assert s.info().rebuilds == r0 + 1
    # fluid moves only in the step's first kernel, before the rebin: keys match the positions now
    x = s.field("pos")
    c = np.floor((x[fl] - lo) / cw).astype(np.int64) % cc
    key = (c[:, 0] * cc[1] + c[:, 1]) * cc[2] + c[:, 2]
    agree = np.mean(key == s.field("key")[fl])
    assert agree == 1.0 if prec == D.PREC_FP64 else agree > 0.999


@pytest.mark.parametrize("prec", [D.PREC_FP64, D.PREC_FP32])
def test_count_nonfinite(prec):
    """dmph_count_nonfinite counts, on the device, the NaN and infinite values
    of a field: none in the state of a normal run (mix3d, 3 steps); exactly the
    three velocity components written as NaN, +inf and -inf through
    dmph_set_field (two of one particle, one of another), while positions stay
    clean; integer fields count 0.  The handle is closed without stepping."""
    g = G.load("mix3d")
    s = _solver(g, precision=prec)
    s.step(3)
    s.sync()
    for f in ("pos", "vel", "pressure_p", "vol_strain_p", "mass"):
        assert s.count_nonfinite(f) == 0, f
    v = s.field("vel")
    v[5, 0], v[5, 2], v[17, 1] = np.nan, np.inf, -np.inf
    s.set_field("vel", v)
    assert s.count_nonfinite("vel") == 3
    assert s.count_nonfinite("pos") == 0
    assert s.count_nonfinite("type") == 0
    s.close()


def test_autotune_keeps_state_and_tolerance():
    """dmph_autotune picks 2, 4, 8 or 16 lanes per sweep without advancing or
    changing the state; the next steps stay within the FP32 tolerances."""
    g = G.load("mix3d")
    s = _solver(g, precision=D.PREC_FP32)
    o = P.Oracle(P.params_from_dict(G.params(g)), G.scene(g), sort_mode=P.SORT_CANONICAL)
    s.step(2)
    o.step(2)
    fields = ("pos", "vel", "force", "pressure_p", "vol_strain_p", "drag_c", "alpha")
    before = {f: s.field(f) for f in fields}
    t0 = s.time()
    lanes = s.autotune(3)
    assert set(lanes) <= {2, 4, 8, 16} and tuple(s.info().lanes) == lanes
    assert set(s.info().block) <= {64, 128, 256}
    assert s.info().wall_lanes in (1, 2, 4)
    assert s.time() == t0
    for f in fields:
        np.testing.assert_array_equal(s.field(f), before[f], err_msg=f)
    s.step(1)
    o.step(1)
    typ = o.field("type")
    fl = typ < 2
    F, Fr, m = s.field("force"), o.field("force"), o.field("mass")
    acc = np.linalg.norm((F[fl] - Fr[fl]) / m[fl, None]) / np.linalg.norm(Fr[fl] / m[fl, None])
    assert acc <= 1e-4


def _row_sets(ptr, idx):
    return [frozenset(idx[ptr[i]:ptr[i + 1]].tolist()) for i in range(len(ptr) - 1)]


@pytest.mark.parametrize("name", ["c1", "mix3d", "st3d", "mix2d", "st2d", "wall2d"])
def test_fp32_rows_hold_the_fp64_pair_sets(name):
    """FP32 rows keep near entries at the front and far ones at the back of the
    row's capacity.  At the first binning their membership is decided in FP64
    on the input positions, so read through dmph_pairs they hold exactly the
    FP64 rows' pairs (canonical order): no pair differs, none is duplicated."""
    sc, prm, L = _case(name)
    s32 = D.Solver(D.Params.from_dict(prm), sc, precision=D.PREC_FP32)
    s64 = D.Solver(D.Params.from_dict(prm), sc, precision=D.PREC_FP64)
    np.testing.assert_array_equal(s32.field("id", D.ORDER_SLOT), s64.field("id", D.ORDER_SLOT))
    p32, i32 = s32.pairs()
    p64, i64 = s64.pairs()
    np.testing.assert_array_equal(np.diff(p32), np.diff(p64))
    r32, r64 = _row_sets(p32, i32), _row_sets(p64, i64)
    assert all(len(r) == n for r, n in zip(r32, np.diff(p32)))   # no duplicates
    differ = sum(len(a ^ b) for a, b in zip(r32, r64))
    assert differ == 0


def _fp64_pairs_of_fp32_state(s, dmin, cut2, keep):
    """The rows an FP64 brute-force test gives on an FP32 handle's state, for
    the slots in `keep` and partners in `keep`: separations
    shift x CellWidth + (o_j - o_i) of the stored FP32 offsets o (recovered
    exactly from the absolute positions and the cell keys), the stencil shift
    of j's cell relative to i's in {-1, 0, 1} per axis, r^2 <= cut2."""
    info = s.info()
    cc, cw = np.array(info.cell_count[:]), np.array(info.cell_width[:])
    key = s.field("key", D.ORDER_SLOT)
    cell = key % info.cells
    co = np.stack([cell // (cc[1] * cc[2]), (cell // cc[2]) % cc[1], cell % cc[2]], axis=1)
    o = ((s.field("pos", D.ORDER_SLOT) - dmin) - co * cw).astype(np.float32).astype(np.float64)
    rows = {}
    for i in np.nonzero(keep)[0]:
        dc = co - co[i]
        dc = np.where(dc > 1, dc - cc, np.where(dc < -1, dc + cc, dc))
        ok = np.all(np.abs(dc) <= 1, axis=1) & keep
        x = dc * cw + (o - o[i])
        r2 = (x * x).sum(axis=1)
        rows[i] = frozenset(np.nonzero(ok & (r2 <= cut2))[0].tolist())
    return rows


@pytest.mark.parametrize("name", ["mix3d", "mix2d", "wall2d"])
def test_fp32_rows_after_steps_hold_the_fp64_pairs_of_the_state(name):
    """After 5 steps with rebinning every step, each FP32 row holds exactly the
    pairs an FP64 brute-force search finds on the FP32 state the handle holds
    (cutoff MaxRadius + 0.1 d, periodic stencil).  Checked for fluid and wall
    particles with fluid and wall partners: these move only before the
    rebinning in a step, solids also after it (the DEM sub-steps).  mix3d's and
    mix2d's walls stand still, so their rebinnings leave the walls out of the
    sort; wall2d starts 3 steps before its walls begin to move (t = 0.2 s), so
    its rebinnings go from that to binning the walls again."""
    sc, prm, L = _case(name)
    if name == "wall2d":
        sc.time = 0.2 - 3 * prm["Dt"]
    s = D.Solver(D.Params.from_dict(prm), sc, precision=D.PREC_FP32, rebuild=D.REBUILD_EVERY_STEP)
    s.step(5)
    s.sync()
    ptr, idx = s.pairs()
    cut = s.info().max_radius + 0.1 * sc.d
    typ = s.field("type", D.ORDER_SLOT)
    keep = (typ < 2) | (typ >= 4)
    got = _row_sets(ptr, idx)
    ref = _fp64_pairs_of_fp32_state(s, np.asarray(sc.domain_min, float), cut * cut, keep)
    differ = sum(len(frozenset(j for j in got[i] if keep[j]) ^ r) for i, r in ref.items())
    assert len(ref) > 100 and differ == 0
    s.close()

