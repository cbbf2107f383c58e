"""The large single scene (C5 channel, 12.9M particles) on the GPU at its full
size.  No oracle runs at this size; the checks are size-independent: every
value finite, walls fixed, each rigid body's shape kept, body masses equal to
the sum of their members, the clock advanced by exactly the steps taken, and
the fluid column starting to fall."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_channel_c5_full_size_keeps_invariants():
    import dem_mph_amd as D
    from dem_mph_amd import scenes as S
    sc = S.channel_c5(seed=1)
    params = S.params_c5()
    print("channel_c5: %d particles" % sc.n, flush=True)
    s = D.Solver(D.Params.from_dict(params), sc, precision=D.PREC_FP32, rigid=True)
    b0 = s.bodies()
    steps = 20
    s.step(steps)
    s.sync()
    assert abs(s.time() - steps * params["Dt"]) < 1e-12
    pos, vel = s.field("pos"), s.field("vel")
    assert np.isfinite(pos).all() and np.isfinite(vel).all()
    wall = sc.type >= S.WALL
    np.testing.assert_allclose(pos[wall], sc.pos[wall], rtol=0, atol=2e-6)   # FP32 cell offsets
    b1 = s.bodies()
    np.testing.assert_array_equal(b1["mass"], b0["mass"])
    d3 = sc.d ** 3
    members = np.bincount(sc.body[sc.body >= 0], weights=sc.density[sc.body >= 0] * d3)
    np.testing.assert_allclose(b0["mass"], members, rtol=1e-12)
    # rigid shapes: member positions relative to the body centre keep their lengths
    rng = np.random.default_rng(0)
    for b in rng.choice(len(members), 16, replace=False):
        m = np.flatnonzero(sc.body == b)
        r0 = np.linalg.norm(sc.pos[m] - b0["x"][b], axis=1)
        r1 = np.linalg.norm(pos[m] - b1["x"][b], axis=1)
        np.testing.assert_allclose(r1, r0, rtol=0, atol=5e-6)
    fluid = sc.type < S.SOLID
    assert vel[fluid, 2].mean() < 0.0
