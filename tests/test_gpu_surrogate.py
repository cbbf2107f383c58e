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
