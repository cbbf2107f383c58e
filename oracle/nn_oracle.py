"""FP64 CPU oracle of the surrogate network (test infrastructure only).

Restates the network of dem_mph_amd/surrogate.py -- encoder, six rounds of edge
and node MLPs on the radius graph, decoder, masked mean squared error -- with
its forward pass and a hand-written backward pass in numpy float64.  The tests
pin the backward pass against central finite differences of the forward pass,
then use it as the reference for the GPU's bf16 MFMA gradients.  The network has
no counterpart in the reference (source/DEM-MPH.cpp has no learned component).

Parameters use the surrogate's state_dict names: enc1, enc2, r{k}_ab, r{k}_c,
r{k}_w2, r{k}_n1, r{k}_n2, dec1, dec2, each with ".w" [n_out, n_in] and, where
the layer has one, ".b".
"""
from __future__ import annotations

from typing import Dict, Tuple

import numpy as np

ROUNDS = 6
W = 128


def relu(x):
    return np.maximum(x, 0.0)


def edges(rp: np.ndarray, ci: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    n = len(rp) - 1
    dst = np.repeat(np.arange(n), np.diff(rp))
    return dst, np.asarray(ci[: rp[-1]], np.int64)


def edge_features(pos: np.ndarray, dst: np.ndarray, src: np.ndarray, radius: float) -> np.ndarray:
    d = (pos[src] - pos[dst]) / radius
    return np.c_[d, np.linalg.norm(pos[src] - pos[dst], axis=1) / radius]


def forward(P: Dict[str, np.ndarray], x0: np.ndarray, pos: np.ndarray, rp: np.ndarray, ci: np.ndarray,
            radius: float):
    """y [N, 3] and the cache for backward()."""
    n = len(x0)
    dst, src = edges(rp, ci)
    eps = edge_features(pos, dst, src, radius)
    c = {"x0": x0, "dst": dst, "src": src, "eps": eps, "rounds": []}
    z1 = x0 @ P["enc1.w"].T + P["enc1.b"]
    h1 = relu(z1)
    h = h1 @ P["enc2.w"].T + P["enc2.b"]
    c["h1"] = h1
    for k in range(ROUNDS):
        pre = "r%d_" % k
        ab = h @ P[pre + "ab.w"].T
        z = ab[dst, :W] + ab[src, W:] + eps @ P[pre + "c.w"].T + P[pre + "c.b"]
        m = relu(z)
        s = np.zeros((n, W))
        np.add.at(s, dst, m)
        agg = s @ P[pre + "w2.w"].T
        hin = np.c_[h, agg]
        u = relu(hin @ P[pre + "n1.w"].T + P[pre + "n1.b"])
        c["rounds"].append({"h": h, "z": z, "s": s, "hin": hin, "u": u})
        h = h + u @ P[pre + "n2.w"].T + P[pre + "n2.b"]
    g = relu(h @ P["dec1.w"].T + P["dec1.b"])
    y = g @ P["dec2.w"].T + P["dec2.b"]
    c["h"] = h
    c["g"] = g
    return y, c


def loss(y: np.ndarray, target: np.ndarray, mask: np.ndarray, count: int) -> float:
    """sum_i mask[i] |y_i - target_i|^2 / count; mask holds per-node weights (0 = left out)."""
    d = y - target
    return float((mask[:, None] * d * d).sum() / count)


def backward(P: Dict[str, np.ndarray], c, y: np.ndarray, target: np.ndarray, mask: np.ndarray,
             count: int) -> Dict[str, np.ndarray]:
    """Gradients of loss() with respect to every parameter."""
    G = {}
    dst, src, eps = c["dst"], c["src"], c["eps"]
    n = len(y)
    dy = 2.0 * (y - target) * mask[:, None] / count
    G["dec2.w"] = dy.T @ c["g"]
    G["dec2.b"] = dy.sum(axis=0)
    dg = (dy @ P["dec2.w"]) * (c["g"] > 0)
    G["dec1.w"] = dg.T @ c["h"]
    G["dec1.b"] = dg.sum(axis=0)
    dh = dg @ P["dec1.w"]
    for k in reversed(range(ROUNDS)):
        pre = "r%d_" % k
        rc = c["rounds"][k]
        G[pre + "n2.w"] = dh.T @ rc["u"]
        G[pre + "n2.b"] = dh.sum(axis=0)
        dz = (dh @ P[pre + "n2.w"]) * (rc["u"] > 0)
        G[pre + "n1.w"] = dz.T @ rc["hin"]
        G[pre + "n1.b"] = dz.sum(axis=0)
        dhin = dz @ P[pre + "n1.w"]
        dh = dh + dhin[:, :W]
        dagg = dhin[:, W:]
        G[pre + "w2.w"] = dagg.T @ rc["s"]
        ds = dagg @ P[pre + "w2.w"]
        dm = ds[dst] * (rc["z"] > 0)
        G[pre + "c.w"] = dm.T @ eps
        G[pre + "c.b"] = dm.sum(axis=0)
        da = np.zeros((n, W))
        db = np.zeros((n, W))
        np.add.at(da, dst, dm)
        np.add.at(db, src, dm)
        dab = np.c_[da, db]
        G[pre + "ab.w"] = dab.T @ rc["h"]
        dh = dh + dab @ P[pre + "ab.w"]
    G["enc2.w"] = dh.T @ c["h1"]
    G["enc2.b"] = dh.sum(axis=0)
    dz1 = (dh @ P["enc2.w"]) * (c["h1"] > 0)
    G["enc1.w"] = dz1.T @ c["x0"]
    G["enc1.b"] = dz1.sum(axis=0)
    return G


def radius_graph(pos: np.ndarray, radius: float) -> Tuple[np.ndarray, np.ndarray]:
    """Brute-force CSR of all pairs within radius (self excluded), rows ascending by node."""
    d2 = ((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1)
    adj = (d2 <= radius * radius) & ~np.eye(len(pos), dtype=bool)
    rp = np.r_[0, np.cumsum(adj.sum(axis=1))]
    return rp.astype(np.int64), np.nonzero(adj)[1].astype(np.int64)


def random_params(seed: int, scale: float = 1.0) -> Dict[str, np.ndarray]:
    rng = np.random.default_rng(seed)
    P = {}

    def lin(name, n_out, n_in, bias=True, s=1.0):
        P[name + ".w"] = rng.standard_normal((n_out, n_in)) * np.sqrt(2.0 / n_in) * s * scale
        if bias:
            P[name + ".b"] = rng.standard_normal(n_out) * 0.1 * scale

    lin("enc1", W, 32)
    lin("enc2", W, W, s=0.5)
    for k in range(ROUNDS):
        pre = "r%d_" % k
        lin(pre + "ab", 2 * W, W, bias=False, s=0.5)
        lin(pre + "c", W, 4)
        lin(pre + "w2", W, W, bias=False, s=0.3)
        lin(pre + "n1", W, 2 * W)
        lin(pre + "n2", W, W, s=0.3)
    lin("dec1", W, W)
    lin("dec2", 3, W, s=0.3)
    return P
