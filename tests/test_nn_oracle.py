"""The surrogate's FP64 oracle (CPU): its backward pass against central finite
differences of its forward pass, and its brute-force radius graph against the
definition."""
import numpy as np
import pytest

import nn_oracle as NO


def _problem(seed=3, n=40):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, 0.1, (n, 3))
    radius = 0.035
    rp, ci = NO.radius_graph(pos, radius)
    x0 = rng.standard_normal((n, 32))
    target = rng.standard_normal((n, 3))
    # integer per-node loss weights as training uses them: 0 (left out), 1, 5
    mask = rng.choice([0.0, 1.0, 5.0], size=n, p=[0.3, 0.5, 0.2])
    return pos, radius, rp, ci, x0, target, mask, int(mask.sum())


def test_radius_graph_is_symmetric_and_exact():
    pos, radius, rp, ci = _problem()[:4]
    dst, src = NO.edges(rp, ci)
    assert len(dst) > 3 * len(pos)   # a few neighbours each
    pairs = set(zip(dst.tolist(), src.tolist()))
    assert all((j, i) in pairs for i, j in pairs)
    d = np.linalg.norm(pos[dst] - pos[src], axis=1)
    assert (d <= radius).all() and (dst != src).all()
    n_in = sum(1 for i in range(len(pos)) for j in range(len(pos))
               if i != j and np.linalg.norm(pos[i] - pos[j]) <= radius)
    assert n_in == len(dst)


@pytest.mark.parametrize("name", ["enc1.w", "enc1.b", "enc2.w", "r0_ab.w", "r0_c.w", "r0_c.b", "r2_w2.w",
                                  "r3_n1.w", "r3_n1.b", "r5_n2.w", "r5_n2.b", "dec1.w", "dec2.w", "dec2.b"])
def test_backward_matches_finite_differences(name):
    pos, radius, rp, ci, x0, target, mask, count = _problem()
    P = NO.random_params(7)
    y, c = NO.forward(P, x0, pos, rp, ci, radius)
    G = NO.backward(P, c, y, target, mask, count)
    rng = np.random.default_rng(sum(map(ord, name)))
    idx = [tuple(rng.integers(0, s) for s in P[name].shape) for _ in range(4)]
    h = 1e-6
    for ix in idx:
        old = P[name][ix]
        P[name][ix] = old + h
        lp = NO.loss(NO.forward(P, x0, pos, rp, ci, radius)[0], target, mask, count)
        P[name][ix] = old - h
        lm = NO.loss(NO.forward(P, x0, pos, rp, ci, radius)[0], target, mask, count)
        P[name][ix] = old
        fd = (lp - lm) / (2 * h)
        assert abs(G[name][ix] - fd) <= 1e-6 * max(1.0, abs(fd)), (name, ix, G[name][ix], fd)


def test_every_parameter_gets_a_gradient():
    pos, radius, rp, ci, x0, target, mask, count = _problem()
    P = NO.random_params(7)
    y, c = NO.forward(P, x0, pos, rp, ci, radius)
    G = NO.backward(P, c, y, target, mask, count)
    assert set(G) == set(P)
    for k in P:
        assert G[k].shape == P[k].shape
        assert np.abs(G[k]).max() > 0.0, k
