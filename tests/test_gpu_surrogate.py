"""The trained surrogate (dem_mph_amd/surrogate_c3.npz) on held-out C3-style
scenes: what the network gains over the same rollout without it."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dem_mph_amd",
                       "surrogate_c3.npz")


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no trained surrogate weights in the tree")
def test_trained_surrogate_beats_rollouts_without_the_network():
    """Sixteen held-out scenes (a seed neither training nor model selection
    used), 100 surrogate steps (1000 solver steps, 0.05 s) from t = 0.1 s,
    re-anchored every 200 solver steps: debris-centroid RMS error against the
    plain solver is below both baselines -- constant velocity and gravity only
    -- and no more scenes exceed the 1 %-of-tank tolerance."""
    from dem_mph_amd import scenes as S
    from dem_mph_amd import surrogate as SG
    sur = SG.load(WEIGHTS)
    ev = S.ensemble_c3(16, seed=7777)
    r = SG.evaluate(sur, ev, S.params_c3(), 2000, 100, 200, 0.008)
    net, still, grav = r["surrogate"], r["still"], r["gravity"]
    print("\n[surrogate-gain] net %.2e still %.2e gravity %.2e (debris RMS, m); fell back %d/%d/%d"
          % (net["debris_centroid_rms_m_mean"], still["debris_centroid_rms_m_mean"],
             grav["debris_centroid_rms_m_mean"], net["fell_back"], still["fell_back"], grav["fell_back"]))
    assert np.isfinite(net["debris_centroid_rms_m_mean"]) and np.isfinite(net["fluid_rms_m_mean"])
    assert net["debris_centroid_rms_m_mean"] < still["debris_centroid_rms_m_mean"]
    assert net["debris_centroid_rms_m_mean"] < grav["debris_centroid_rms_m_mean"]
    assert net["fell_back"] <= min(still["fell_back"], grav["fell_back"])


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no trained surrogate weights in the tree")
def test_surrogate_with_fallback_runs_as_deployed():
    """dem_mph_amd.surrogate.hybrid, the surrogate with its plain-solver
    fallback from t = 0 (solver warm-up, re-anchoring and the probe judge every
    8 strides, fallen-back scenes on the solver), on four held-out scenes for
    24 strides after the 4-stride warm-up (the probes' strides count towards
    them): the clock reaches 28 strides either way; with an
    infinite tolerance every scene stays on the surrogate, with a negative one
    every scene falls back at the first check: the solver recomputes it from
    its last anchored solver state (the warm-up's), so the run ends where the
    plain solver batch run from t = 0 for the same 280 steps ends (debris
    centroids within 1e-4 m: the fallback batch is rebuilt from FP32 states,
    which rounds them).  Final states finite throughout, every debris body
    inside the domain."""
    from dem_mph_amd import scenes as S
    from dem_mph_amd import surrogate as SG
    import dem_mph_amd as D
    sur = SG.load(WEIGHTS)
    ev = S.ensemble_c3(4, seed=7778)
    params = S.params_c3()
    keep = SG.hybrid(sur, ev, params, 24, reanchor=8, tol=1e9)
    assert keep["strides"] == 4 + 24 and keep["fell_back"] == 0 and keep["on_surrogate_at_end"] == 4
    assert len(keep["judged"]) == 2 and all(not j["fell_back"] for j in keep["judged"])
    fall = SG.hybrid(sur, ev, params, 24, reanchor=8, tol=-1.0)
    assert fall["strides"] == 4 + 24 and fall["fell_back"] == 4 and fall["on_surrogate_at_end"] == 0
    assert len(fall["judged"]) == 1 and fall["judged"][0]["fell_back"] == [0, 1, 2, 3]
    for r in (keep, fall):
        for k in range(4):
            assert np.isfinite(r["final"][k]).all()
            c = SG.debris_centroids(ev[k], r["final"][k])
            assert np.all(c >= ev[k].domain_min) and np.all(c <= ev[k].domain_max)
    s = D.Solver(D.Params.from_dict(params), ev, precision=D.PREC_FP32, compensated=False,
                 rigid_precision=D.PREC_FP32, rigid=True)
    s.step(280)
    P = s.field("pos")
    off = s.scene_offsets
    s.close()
    d = max(np.abs(SG.debris_centroids(ev[k], fall["final"][k]) -
                   SG.debris_centroids(ev[k], P[int(off[k]):int(off[k + 1])])).max() for k in range(4))
    print("\n[hybrid] fallen back vs the plain solver: debris centroids up to %.2e m apart" % d)
    assert d <= 1e-4
