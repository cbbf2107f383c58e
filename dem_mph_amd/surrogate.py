"""Learned surrogate for the DEM-MPH step: a message-passing network on the
radius graph (r <= 3 d) that predicts particle accelerations at a 10 dt stride.

No reference counterpart (the reference has no learned component).  The
network, per node (particle):

    x0  = features: 5 velocities at 10 dt spacing (newest first, scaled),
          class one-hot (fluid, debris, wall), 6 wall-face distances clipped
          to [0, 1] in units of the radius, zero padding to 32
    h   = We2 relu(We1 x0 + be1) + be2                         (encoder, 2 layers)
    6 rounds:
      A | B = Wab h                                            (per node, 256)
      S_i   = sum_{j in N(i)} relu(A_i + B_j + Wc eps_ij + bc)  (edge layer 1)
      agg_i = W2 S_i                                           (edge layer 2)
      h    += Wn2 relu(Wn1 [h | agg] + bn1) + bn2              (node MLP)
    y   = Wd2 relu(Wd1 h + bd1) + bd2                          (decoder, 3 outputs)

The edge MLP is the 2-layer width-128 MLP m_ij = W2 relu(W1 [h_i, h_j, eps_ij]
+ b) with its first layer split by input (W1 = [Wa | Wb | Wc]) and its linear
second layer applied after the sum over j -- the same function, with every
dense product on a per-node bf16 MFMA GEMM and the per-edge work a gather,
add, ReLU and sum.  The backward pass recomputes each edge's ReLU mask from the
per-node projections instead of storing per-edge activations.

Torch tensors are used only as device memory; every operation is a kernel of
libdemmph_nn.so (include/dem_mph_nn.h) called through its C ABI.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional, Tuple

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
NN_LIB = os.path.join(PKG, "libdemmph_nn.so")

WIDTH = 128
ROUNDS = 6
FEAT = 32
GEMM_RELU, GEMM_F16_OUT = 1, 2   # dmph_nn_gemm flags (DMPH_NN_RELU, DMPH_NN_F16_OUT)
HIST = 5


class NnGrid(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("inv_w", C.c_float), ("cc", C.c_int * 3), ("ncell", C.c_int),
                ("nscene", C.c_int)]


class NnFeat(C.Structure):
    _fields_ = [("vscale", C.c_float), ("tlo", C.c_float * 3), ("thi", C.c_float * 3), ("inv_r", C.c_float)]


class NnFold(C.Structure):
    _fields_ = [("pos", C.c_void_p), ("Wt", C.c_void_p), ("origin", C.c_float * 3), ("inv_r", C.c_float)]


# The position fold (dmph_nn_edge_fwd_f16) rounds Wc3 (x - origin) / radius to
# half in the node terms: used while the nodes' box reaches at most this many
# radii from its centre (the C3 tank: 14), else the edge sums take the edge
# features whole.
FOLD_MAX_RADII = 16.0

_nn = None


def nnlib():
    """The surrogate kernels; a missing library is an error (there is no fallback)."""
    global _nn
    if _nn is None:
        if not os.path.exists(NN_LIB):
            raise RuntimeError("%s is not built (python -m dem_mph_amd.build)" % NN_LIB)
        from . import share_hip_runtime
        share_hip_runtime()
        _nn = C.CDLL(NN_LIB)
        for f in ("dmph_nn_dw_scratch", "dmph_nn_edge_scratch", "dmph_nn_head_scratch", "dmph_nn_mse_scratch",
                  "dmph_nn_graph_scratch", "dmph_nn_bbox_scratch"):
            getattr(_nn, f).restype = C.c_size_t
    return _nn


def _p(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _chk(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError("%s failed (%d)" % (what, rc))


class Layer:
    """FP32 master weight [n_out][n_in] (+ bias), bf16 copy and transposed bf16
    copy for the GEMMs, gradient and Adam moments."""

    def __init__(self, torch, dev, n_out: int, n_in: int, bias: bool, scale: float, rng: np.random.Generator,
                 bf16: bool = True):
        w = rng.standard_normal((n_out, n_in)) * scale
        self.w = torch.tensor(w, dtype=torch.float32, device=dev)
        self.b = torch.zeros(n_out, dtype=torch.float32, device=dev) if bias else None
        self.gw = torch.zeros_like(self.w)
        self.gb = torch.zeros_like(self.b) if bias else None
        self.mw, self.vw = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.mb = torch.zeros_like(self.b) if bias else None
        self.vb = torch.zeros_like(self.b) if bias else None
        self.n_out, self.n_in = n_out, n_in
        self.wb = torch.zeros((n_out, n_in), dtype=torch.int16, device=dev) if bf16 else None
        self.wbt = torch.zeros((n_in, n_out), dtype=torch.int16, device=dev) if bf16 else None

    def tensors(self) -> Dict[str, object]:
        d = {"w": self.w}
        if self.b is not None:
            d["b"] = self.b
        return d


class Surrogate:
    """The network, its parameters and the training/inference passes on one GPU."""

    def __init__(self, d: float, tank_lo, tank_hi, vscale: float, ascale: float, seed: int = 0, kstep: int = 32,
                 device: str = "cuda:0"):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.d = d
        self.radius = 3.0 * d
        self.tank_lo = np.asarray(tank_lo, float)
        self.tank_hi = np.asarray(tank_hi, float)
        self.vscale = float(vscale)     # features: v * vscale
        self.ascale = float(ascale)     # targets: a * ascale
        self.kstep = kstep
        rng = np.random.default_rng(seed)
        T, W = torch, WIDTH
        he = lambda n_in: np.sqrt(2.0 / n_in)
        self.enc1 = Layer(T, self.dev, W, FEAT, True, he(FEAT), rng)
        self.enc2 = Layer(T, self.dev, W, W, True, he(W) * 0.5, rng)
        self.rounds = []
        for _ in range(ROUNDS):
            r = {
                "ab": Layer(T, self.dev, 2 * W, W, False, he(W) * 0.5, rng),
                "c": Layer(T, self.dev, W, 4, True, 0.5, rng, bf16=False),
                "w2": Layer(T, self.dev, W, W, False, he(W) * 0.1, rng),
                "n1": Layer(T, self.dev, W, 2 * W, True, he(2 * W), rng),
                "n2": Layer(T, self.dev, W, W, True, he(W) * 0.1, rng),
            }
            self.rounds.append(r)
        self.dec1 = Layer(T, self.dev, W, W, True, he(W), rng)
        self.dec2 = Layer(T, self.dev, 3, W, True, he(W) * 0.1, rng, bf16=False)
        self.layers: List[Layer] = [self.enc1, self.enc2] + [l for r in self.rounds for l in r.values()] + \
                                   [self.dec1, self.dec2]
        self.adam_step = 0
        self._refresh_bf16()
        self.stream = C.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        self._ws = None
        self._graph_ws = None
        self.graph_eps = None   # edge features of the last graph(..., eps=True)
        self._bbox = None
        self._box = None        # host copy of the last graph's bounding box, with the positions it belongs to
        self._folded = None   # inference: per round [n1 left | n1 right @ w2] (rebuilt after a weight change)
        self._wt = None       # inference: per round Wc[:, 0:3] transposed (the position fold)
        self.rows_dist = False  # the last graph_rows wrote distances only (for the fold)
        self.pos_fold = True  # fold the edge features' linear part into the node terms where the box allows

    def at_spacing(self, d: float) -> "Surrogate":
        """Use the network on particles of spacing d: the graph radius becomes
        3 d.  Every input is in units of the radius or of the velocity scale,
        so the shapes fit; the accuracy at a spacing other than the training
        one is not established (used for throughput on other scenes)."""
        self.d = d
        self.radius = 3.0 * d
        self._bbox = None
        return self

    def in_tank(self, lo, hi) -> "Surrogate":
        """Use the network in a tank whose interior is the box [lo, hi]: the
        clipped wall distances of the node features are taken from its faces."""
        self.tank_lo = np.asarray(lo, float)
        self.tank_hi = np.asarray(hi, float)
        return self

    # ------------------------------------------------------------------ parameters
    def _refresh_bf16(self) -> None:
        """bf16 copies of every GEMM weight (an Adam step with lr 0 rewrites them)."""
        lib = nnlib()
        s = C.c_void_p(self.torch.cuda.current_stream(self.dev).cuda_stream)
        for l in self.layers:
            if l.wb is None:
                continue
            z = self.torch.zeros_like(l.w)
            _chk(lib.dmph_nn_adam(_p(l.w), _p(z), _p(self.torch.zeros_like(l.w)), _p(self.torch.zeros_like(l.w)),
                                  l.w.numel(), C.c_float(0.0), C.c_float(0.9), C.c_float(0.999), C.c_float(1.0), 1,
                                  C.c_float(0.0), _p(l.wb), _p(l.wbt), l.n_out, l.n_in, s), "bf16 copy")

    def state_dict(self) -> Dict[str, np.ndarray]:
        out = {}
        names = ["enc1", "enc2"] + ["r%d_%s" % (k, n) for k in range(ROUNDS) for n in ("ab", "c", "w2", "n1", "n2")] + \
                ["dec1", "dec2"]
        for name, l in zip(names, self.layers):
            out[name + ".w"] = l.w.detach().cpu().numpy().astype(np.float64)
            if l.b is not None:
                out[name + ".b"] = l.b.detach().cpu().numpy().astype(np.float64)
        return out

    def grads(self) -> Dict[str, np.ndarray]:
        out = {}
        names = ["enc1", "enc2"] + ["r%d_%s" % (k, n) for k in range(ROUNDS) for n in ("ab", "c", "w2", "n1", "n2")] + \
                ["dec1", "dec2"]
        for name, l in zip(names, self.layers):
            out[name + ".w"] = l.gw.detach().cpu().numpy().astype(np.float64)
            if l.gb is not None:
                out[name + ".b"] = l.gb.detach().cpu().numpy().astype(np.float64)
        return out

    def load_state(self, sd: Dict[str, np.ndarray]) -> None:
        names = ["enc1", "enc2"] + ["r%d_%s" % (k, n) for k in range(ROUNDS) for n in ("ab", "c", "w2", "n1", "n2")] + \
                ["dec1", "dec2"]
        for name, l in zip(names, self.layers):
            l.w.copy_(self.torch.tensor(sd[name + ".w"], dtype=self.torch.float32))
            if l.b is not None:
                l.b.copy_(self.torch.tensor(sd[name + ".b"], dtype=self.torch.float32))
        self._refresh_bf16()
        self._folded = None
        self._wt = None

    # ------------------------------------------------------------------ graph
    def graph(self, pos, scene=None, nscene: int = 1, eps: bool = False) -> Tuple[object, object, int]:
        """CSR radius graph (r <= 3 d) of the nodes at pos [N, 3] (FP32, device);
        with `eps` the edge features for inference are written alongside
        (self.graph_eps, for forward(..., eps=...))."""
        torch = self.torch
        N = pos.shape[0]
        lib = nnlib()
        stream = C.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        g, nk = self._grid(pos, nscene, stream)
        ws = self._graph_work(N, nk)
        rp = torch.empty(N + 1, dtype=torch.int32, device=self.dev)
        nnz = C.c_int(0)
        for _ in range(2):
            ci = ws["ci"]
            if eps and ws["eps"] is None and ci is not None:
                ws["eps"] = torch.empty(2 * ws["cap"], dtype=torch.int32, device=self.dev)
            rc = lib.dmph_nn_graph(_p(pos), _p(scene), N, C.byref(g), C.c_float(self.radius), _p(ws["key"]),
                                   _p(ws["counts"]), _p(ws["start"]), _p(ws["fill"]), _p(ws["order"]),
                                   _p(ws["sorted"]), _p(ws["deg"]), _p(rp), _p(ci), _p(ws["eps"] if eps else None),
                                   ws["cap"], C.byref(nnz), _p(ws["scan"]), stream)
            if rc == -3:
                raise RuntimeError("radius graph: more than 2^31 edges (radius %.3g m for %d nodes)" % (self.radius, N))
            if rc == -2:
                ws["cap"] = min(int(nnz.value * 1.25) + 1024, 2 ** 31 - 1)
                ws["ci"] = torch.empty(ws["cap"], dtype=torch.int32, device=self.dev)
                ws["eps"] = None
                continue
            _chk(rc, "dmph_nn_graph")
            break
        self.graph_eps = ws["eps"] if eps else None
        return rp, ws["ci"], int(nnz.value)

    def _grid(self, pos, nscene: int, stream):
        """Uniform grid over the nodes' bounding box, cells at least the radius wide."""
        torch = self.torch
        N = pos.shape[0]
        lib = nnlib()
        if self._bbox is None:
            self._bbox = (torch.empty(6, dtype=torch.float32, device=self.dev),
                          torch.empty(lib.dmph_nn_bbox_scratch(N) // 4 + 1, dtype=torch.float32, device=self.dev))
        _chk(lib.dmph_nn_bbox(_p(pos), N, _p(self._bbox[0]), _p(self._bbox[1]), stream), "dmph_nn_bbox")
        box = self._bbox[0].cpu().numpy().astype(np.float64)
        self._box = (box, pos.data_ptr(), N)
        lo, hi = box[:3] - 1e-6, box[3:] + 1e-6
        w = self.radius * (1.0 + 1e-5)            # cubic cells at least the radius wide
        cc = np.maximum(np.ceil((hi - lo) / w).astype(int), 1)
        g = NnGrid()
        for k in range(3):
            g.lo[k] = float(lo[k])
            g.cc[k] = int(cc[k])
        g.inv_w = float(1.0 / w)
        g.ncell = int(cc.prod())
        g.nscene = nscene
        return g, g.ncell * nscene

    def _graph_work(self, N: int, nk: int):
        ws = self._graph_ws
        if ws is None or ws["N"] < N or ws["nk"] < nk:
            torch = self.torch
            i32 = lambda n: torch.empty(max(n, 1), dtype=torch.int32, device=self.dev)
            ws = {"N": N, "nk": nk, "key": i32(N), "order": i32(N), "sorted": i32(4 * N), "deg": i32(N),
                  "counts": i32(nk + 1), "start": i32(nk + 1), "fill": i32(nk + 1), "cap": 0, "ci": None, "eps": None,
                  "scan": i32(nnlib().dmph_nn_graph_scratch(N, nk) // 4 + 1)}
            self._graph_ws = ws
        return ws

    def graph_rows(self, pos, scene=None, nscene: int = 1, row_cap: int = 160, dist: Optional[bool] = None):
        """Inference graph in one pass: rows of a fixed capacity (row i at
        i * row_cap, deg[i] entries) with the edge features; the capacity grows
        while rows are cut.  Returns (ci, deg, row_cap, eps), or None where N x
        row_cap passes int32 offsets (then use graph()).  With `dist` the
        features are the distances alone (one half per entry, what the position
        fold reads); None: the distances where the fold applies to these nodes.
        self.rows_dist records which was written."""
        torch = self.torch
        N = pos.shape[0]
        lib = nnlib()
        stream = C.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        g, nk = self._grid(pos, nscene, stream)
        if dist is None:
            dist = self._fold_args(pos, N) is not None
        ws = self._graph_work(N, nk)
        cap = max(ws.get("row_cap", 0), row_cap)
        over = C.c_int(0)
        while True:
            if N * cap > 2 ** 31 - 1:
                return None
            if ws.get("row_cap", 0) != cap or ws.get("rows_n", 0) < N:
                ws["rows_ci"] = torch.empty(N * cap, dtype=torch.int32, device=self.dev)
                ws["rows_eps"] = torch.empty(2 * N * cap, dtype=torch.int32, device=self.dev)
                ws["row_cap"], ws["rows_n"] = cap, N
            rc = lib.dmph_nn_graph_rows(_p(pos), _p(scene), N, C.byref(g), C.c_float(self.radius), _p(ws["key"]),
                                        _p(ws["counts"]), _p(ws["start"]), _p(ws["fill"]), _p(ws["order"]),
                                        _p(ws["sorted"]), _p(ws["deg"]), _p(ws["rows_ci"]),
                                        _p(None if dist else ws["rows_eps"]), _p(ws["rows_eps"] if dist else None), cap,
                                        C.byref(over), _p(ws["scan"]), stream)
            if rc == -2:
                cap = (cap * 3 // 2 + 15) // 16 * 16
                continue
            _chk(rc, "dmph_nn_graph_rows")
            self.rows_dist = bool(dist)
            return ws["rows_ci"], ws["deg"], cap, ws["rows_eps"]

    # ------------------------------------------------------------------ forward / backward
    def _workspace(self, N: int, train: bool = True):
        """Activations for N nodes: every round's (training) or two alternating
        slots (inference)."""
        torch = self.torch
        ws = self._ws
        if ws is not None and ws["N"] >= N and (ws["train"] or not train):
            return ws
        self._ws = None
        bf = lambda *s: torch.zeros(s, dtype=torch.int16, device=self.dev)
        f32 = lambda *s: torch.zeros(s, dtype=torch.float32, device=self.dev)
        W = WIDTH
        nr = ROUNDS if train else 1
        ws = {"N": N, "train": train, "x0": bf(N, FEAT), "h1": bf(N, W),
              "hin": [bf(N, 2 * W) for _ in range(nr + 1)],   # [h_r | agg_r]
              "ab": [bf(N, 2 * W) for _ in range(nr)], "s": [bf(N, W) for _ in range(nr)],
              "u": [bf(N, W) for _ in range(nr)], "g": bf(N, W), "y": f32(N, 3), "loss": f32(1)}
        if train:
            ws.update({"dy": f32(N, 3), "dh": f32(N, W), "dhb": bf(N, W), "t1": bf(N, W), "t2": bf(N, 2 * W),
                       "dab": bf(N, 2 * W), "ds": f32(N, W)})
        groups = 256
        lib = nnlib()
        scr = max(lib.dmph_nn_dw_scratch(N, 256, 256, groups), lib.dmph_nn_edge_scratch(N),
                  lib.dmph_nn_head_scratch(N), lib.dmph_nn_mse_scratch(N))
        ws["scratch"] = torch.zeros(scr // 4 + 1, dtype=torch.float32, device=self.dev)
        ws["groups"] = groups
        self._ws = ws
        return ws

    def features(self, vel_hist, cls, pos, train: bool = True):
        """x0 [N, 32] bf16 from vel_hist [5, N, 3] (newest first), class ids, positions."""
        N = pos.shape[0]
        ws = self._workspace(N, train)
        _chk(nnlib().dmph_nn_features(_p(vel_hist), N, _p(cls), _p(pos), C.byref(self._feat()), _p(ws["x0"]),
                                      self._s()), "features")
        return ws["x0"]

    def _feat(self) -> "NnFeat":
        ft = NnFeat()
        ft.vscale = self.vscale
        for k in range(3):
            ft.tlo[k] = float(self.tank_lo[k])
            ft.thi[k] = float(self.tank_hi[k])
        ft.inv_r = 1.0 / self.radius
        return ft

    def _s(self) -> C.c_void_p:
        return C.c_void_p(self.torch.cuda.current_stream(self.dev).cuda_stream)

    def _gemm(self, X, ldx, layer: Layer, Y, ldy, M, relu, R=None, ldr=0, transposed=False, K=None, Nout=None,
              half=False):
        w = layer.wbt if transposed else layer.wb
        k = K if K is not None else (layer.n_out if transposed else layer.n_in)
        n = Nout if Nout is not None else (layer.n_in if transposed else layer.n_out)
        bias = None if transposed else layer.b
        flags = (GEMM_RELU if relu else 0) | (GEMM_F16_OUT if half else 0)
        _chk(nnlib().dmph_nn_gemm(_p(X), ldx, _p(w), _p(bias), _p(R), ldr, _p(Y), ldy, M, k, n, flags,
                                  self.kstep, self._s()), "gemm")

    def _dw(self, dY, lddy, X, ldx, layer: Layer, M):
        ws = self._ws
        _chk(nnlib().dmph_nn_gemm_dw(_p(dY), lddy, _p(X), ldx, _p(layer.gw), _p(layer.gb), M, layer.n_in, layer.n_out,
                                     ws["groups"], _p(ws["scratch"]), 0, self._s()), "gemm_dw")

    def _fold(self):
        """Inference weights: the aggregation's linear layer folded into the node
        MLP's first layer, Wn1 [h | W2 S] = [Wn1_h | Wn1_a W2] [h | S] (FP32
        product, then the bf16 copy), so inference skips one GEMM per round."""
        if self._folded is None:
            lib = nnlib()
            W = WIDTH
            rng = np.random.default_rng(0)
            out = []
            for L in self.rounds:
                f = Layer(self.torch, self.dev, W, 2 * W, True, 0.0, rng)
                n1, w2 = L["n1"].w.detach().cpu().numpy().astype(np.float64), L["w2"].w.detach().cpu().numpy().astype(np.float64)
                fw = np.concatenate([n1[:, :W], n1[:, W:] @ w2], axis=1)
                f.w.copy_(self.torch.tensor(fw, dtype=self.torch.float32))
                f.b.copy_(L["n1"].b)
                z = self.torch.zeros_like(f.w)
                _chk(lib.dmph_nn_adam(_p(f.w), _p(z), _p(self.torch.zeros_like(f.w)), _p(self.torch.zeros_like(f.w)),
                                      f.w.numel(), C.c_float(0.0), C.c_float(0.9), C.c_float(0.999), C.c_float(1.0), 1,
                                      C.c_float(0.0), _p(f.wb), _p(f.wbt), f.n_out, f.n_in, self._s()), "bf16 copy")
                out.append(f)
            self._folded = out
            self._wt = [L["c"].w[:, :3].t().contiguous() for L in self.rounds]
        return self._folded

    def _fold_args(self, pos, N: int):
        """NnFold for these positions (origin at the centre of their box), or
        None where the box is unknown or too wide for the fold's half rounding."""
        if not self.pos_fold or self._box is None:
            return None
        box, ptr, n = self._box
        if ptr != pos.data_ptr() or n != N:
            return None
        lo, hi = box[:3], box[3:]
        if float(np.max(hi - lo)) * 0.5 / self.radius > FOLD_MAX_RADII:
            return None
        f = NnFold()
        f.pos = pos.data_ptr()
        for k in range(3):
            f.origin[k] = float(0.5 * (lo[k] + hi[k]))
        f.inv_r = 1.0 / self.radius
        return f

    def _fold_for(self, f, r: int):
        """f with the transposed Wc[:, 0:3] of round r (None stays None)."""
        if f is None:
            return None
        f.Wt = self._wt[r].data_ptr()
        return C.byref(f)

    def forward(self, x0, pos, rp, ci, N: int, train: bool = True, eps=None, rows=None, feats=None, order=None):
        """Accelerations (scaled) y [N, 3] FP32; with `train` every round's
        activations are kept for backward, else two slots alternate and the
        aggregation's linear layer is folded into the node MLP (_fold).
        Inference with `feats` (vel_hist, cls) runs the fused kernels (encoder,
        node updates, position fold where the box allows); `order` (a
        permutation, e.g. the graph's cell order) is the order the edge sums
        visit the nodes in."""
        lib = nnlib()
        ws = self._workspace(N, train)
        W = WIDTH
        s = self._s()
        slot = (lambda r: r) if train else (lambda r: 0)
        hslot = (lambda r: r) if train else (lambda r: r % 2)
        fused = not train and self.kstep == 32
        folded = None if train else self._fold()
        # the position fold needs every [A | B] from a fused kernel (encoder included)
        # and rows holding the distances (graph_rows with dist)
        dist_rows = rows is not None and self.rows_dist
        fold = self._fold_args(pos, N) if fused and feats is not None and dist_rows else None
        if dist_rows and fold is None:
            raise RuntimeError("the rows hold distances only: they are summed by the folded inference path")
        if fused and feats is not None:   # features, encoder and the first [A | B] in one kernel
            vel_hist, cls = feats
            _chk(lib.dmph_nn_encode(_p(vel_hist), N, _p(cls), _p(pos), C.byref(self._feat()), _p(self.enc1.wb),
                                    _p(self.enc1.b), _p(self.enc2.wb), _p(self.enc2.b), _p(self.rounds[0]["ab"].wb),
                                    _p(ws["hin"][0]), 2 * W, _p(ws["ab"][0]), self._fold_for(fold, 0), s), "encode")
        else:
            self._gemm(x0, FEAT, self.enc1, ws["h1"], W, N, True)
            self._gemm(ws["h1"], W, self.enc2, ws["hin"][0], 2 * W, N, False)
        if not train and eps is None:   # edge features once per graph, for the half-precision edge sums
            if ws.get("eps") is None or ws["eps"].numel() < 2 * ci.numel():
                ws["eps"] = self.torch.empty(2 * ci.numel(), dtype=self.torch.int32, device=self.dev)
            _chk(lib.dmph_nn_edge_eps(_p(pos), _p(rp), _p(ci), _p(ws["eps"]), N, C.c_float(self.radius), s),
                 "edge_eps")
            eps = ws["eps"]
        for r, L in enumerate(self.rounds):
            hin = ws["hin"][hslot(r)]
            ab, sr, u = ws["ab"][slot(r)], ws["s"][slot(r)], ws["u"][slot(r)]
            if not (fused and (r > 0 or feats is not None)):   # else a fused kernel wrote this round's [A | B]
                self._gemm(hin, 2 * W, L["ab"], ab, 2 * W, N, False, half=not train)
            if fused and r + 1 < ROUNDS:
                _chk(lib.dmph_nn_edge_fwd_f16(_p(ab), _p(rp), _p(ci), _p(eps), _p(L["c"].w), _p(L["c"].b),
                                              _p(hin[:, W:]), 2 * W, N, _p(rows[0] if rows else None),
                                              rows[1] if rows else 0, _p(order), self._fold_for(fold, r), s),
                     "edge_fwd_f16")
                nxt = self.rounds[r + 1]["ab"]
                _chk(lib.dmph_nn_node(_p(hin), 2 * W, _p(folded[r].wb), _p(folded[r].b), _p(L["n2"].wb),
                                      _p(L["n2"].b), _p(nxt.wb), _p(ws["hin"][hslot(r + 1)]), 2 * W, _p(ab), N,
                                      self._fold_for(fold, r + 1), s), "node")
                continue
            if fused:   # the last round and the decoder in one kernel
                _chk(lib.dmph_nn_edge_fwd_f16(_p(ab), _p(rp), _p(ci), _p(eps), _p(L["c"].w), _p(L["c"].b),
                                              _p(hin[:, W:]), 2 * W, N, _p(rows[0] if rows else None),
                                              rows[1] if rows else 0, _p(order), self._fold_for(fold, r), s),
                     "edge_fwd_f16")
                _chk(lib.dmph_nn_node_head(_p(hin), 2 * W, _p(folded[r].wb), _p(folded[r].b), _p(L["n2"].wb),
                                           _p(L["n2"].b), _p(self.dec1.wb), _p(self.dec1.b), _p(self.dec2.w),
                                           _p(self.dec2.b), _p(ws["y"]), N, s), "node_head")
                return ws["y"][:N]
            if train:
                _chk(lib.dmph_nn_edge_fwd(_p(ab), _p(pos), _p(rp), _p(ci), _p(L["c"].w), _p(L["c"].b),
                                          _p(sr), W, N, C.c_float(self.radius), s), "edge_fwd")
                # agg into the right half of hin_r
                self._gemm(sr, W, L["w2"], hin[:, W:], 2 * W, N, False)
                self._gemm(hin, 2 * W, L["n1"], u, W, N, True)
            else:   # S straight into the right half of hin_r, then the folded node layer
                deg, cap = rows if rows is not None else (None, 0)
                _chk(lib.dmph_nn_edge_fwd_f16(_p(ab), _p(rp), _p(ci), _p(eps), _p(L["c"].w), _p(L["c"].b),
                                              _p(hin[:, W:]), 2 * W, N, _p(deg), cap, _p(order), None, s),
                     "edge_fwd_f16")
                self._gemm(hin, 2 * W, folded[r], u, W, N, True)
            self._gemm(u, W, L["n2"], ws["hin"][hslot(r + 1)], 2 * W, N, False, R=hin, ldr=2 * W)
        if train:
            self._gemm(ws["hin"][hslot(ROUNDS)], 2 * W, self.dec1, ws["g"], W, N, True)
            _chk(lib.dmph_nn_head(_p(ws["g"]), _p(self.dec2.w), _p(self.dec2.b), _p(ws["y"]), N, s), "head")
        else:   # the hidden layer never leaves the registers
            _chk(lib.dmph_nn_gemm_head(_p(ws["hin"][hslot(ROUNDS)]), 2 * W, _p(self.dec1.wb), _p(self.dec1.b),
                                       _p(self.dec2.w), _p(self.dec2.b), _p(ws["y"]), N, self.kstep, s), "gemm_head")
        return ws["y"][:N]

    def loss_and_backward(self, x0, pos, rp, ci, N: int, target, mask, count: int, clip: float = 1e30) -> float:
        """MSE of forward() against target [N, 3] (clipped to +-clip) on mask;
        fills every layer's gradient."""
        lib = nnlib()
        ws = self._ws
        W = WIDTH
        s = self._s()
        _chk(lib.dmph_nn_mse(_p(ws["y"]), _p(target), _p(mask), N, count, C.c_float(clip), _p(ws["dy"]),
                             _p(ws["loss"]), _p(ws["scratch"]), s), "mse")
        # decoder
        _chk(lib.dmph_nn_head_bwd(_p(ws["g"]), _p(self.dec2.w), _p(ws["dy"]), _p(ws["t1"]), _p(self.dec2.gw),
                                  _p(self.dec2.gb), N, _p(ws["scratch"]), s), "head_bwd")
        h = ws["hin"][ROUNDS]
        self._dw(ws["t1"], W, h, 2 * W, self.dec1, N)
        self._gemm(ws["t1"], W, self.dec1, ws["dhb"], W, N, False, transposed=True)
        _chk(lib.dmph_nn_bf2f(_p(ws["dhb"]), W, _p(ws["dh"]), W, N, W, 0, s), "bf2f")
        for r in reversed(range(ROUNDS)):
            L = self.rounds[r]
            hin = ws["hin"][r]
            # h_{r+1} = h_r + n2(u_r): dn2 = dh
            _chk(lib.dmph_nn_f2bf(_p(ws["dh"]), W, _p(ws["dhb"]), W, N, W, s), "f2bf")
            self._dw(ws["dhb"], W, ws["u"][r], W, L["n2"], N)
            self._gemm(ws["dhb"], W, L["n2"], ws["t1"], W, N, False, transposed=True)
            _chk(lib.dmph_nn_relu_bwd(_p(ws["u"][r]), W, _p(ws["t1"]), W, _p(ws["t1"]), W, N, W, s), "relu_bwd")
            self._dw(ws["t1"], W, hin, 2 * W, L["n1"], N)
            self._gemm(ws["t1"], W, L["n1"], ws["t2"], 2 * W, N, False, transposed=True)   # [dh | dagg]
            _chk(lib.dmph_nn_bf2f(_p(ws["t2"]), 2 * W, _p(ws["dh"]), W, N, W, 1, s), "bf2f")
            dagg = ws["t2"][:, W:]
            self._dw(dagg, 2 * W, ws["s"][r], W, L["w2"], N)
            self._gemm(dagg, 2 * W, L["w2"], ws["t1"], W, N, False, transposed=True)        # dS (bf16)
            _chk(lib.dmph_nn_bf2f(_p(ws["t1"]), W, _p(ws["ds"]), W, N, W, 0, s), "bf2f")
            _chk(lib.dmph_nn_edge_bwd(_p(ws["ab"][r]), _p(pos), _p(rp), _p(ci), _p(L["c"].w), _p(L["c"].b),
                                      _p(ws["ds"]), _p(ws["dab"]), _p(L["c"].gw), _p(L["c"].gb), N,
                                      C.c_float(self.radius), _p(ws["scratch"]), 0, s), "edge_bwd")
            self._dw(ws["dab"], 2 * W, hin, 2 * W, L["ab"], N)
            self._gemm(ws["dab"], 2 * W, L["ab"], ws["dhb"], W, N, False, transposed=True)
            _chk(lib.dmph_nn_bf2f(_p(ws["dhb"]), W, _p(ws["dh"]), W, N, W, 1, s), "bf2f")
        # encoder: h0 = enc2(h1)
        _chk(lib.dmph_nn_f2bf(_p(ws["dh"]), W, _p(ws["dhb"]), W, N, W, s), "f2bf")
        self._dw(ws["dhb"], W, ws["h1"], W, self.enc2, N)
        self._gemm(ws["dhb"], W, self.enc2, ws["t1"], W, N, False, transposed=True)
        _chk(lib.dmph_nn_relu_bwd(_p(ws["h1"]), W, _p(ws["t1"]), W, _p(ws["t1"]), W, N, W, s), "relu_bwd")
        self._dw(ws["t1"], W, x0, FEAT, self.enc1, N)
        return float(ws["loss"].item())

    def adam(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8, wd: float = 0.0) -> None:
        self.adam_step += 1
        self._folded = None
        lib = nnlib()
        s = self._s()
        for l in self.layers:
            _chk(lib.dmph_nn_adam(_p(l.w), _p(l.gw), _p(l.mw), _p(l.vw), l.w.numel(), C.c_float(lr), C.c_float(b1),
                                  C.c_float(b2), C.c_float(eps), self.adam_step, C.c_float(wd), _p(l.wb), _p(l.wbt),
                                  l.n_out, l.n_in, s), "adam")
            if l.b is not None:
                _chk(lib.dmph_nn_adam(_p(l.b), _p(l.gb), _p(l.mb), _p(l.vb), l.b.numel(), C.c_float(lr),
                                      C.c_float(b1), C.c_float(b2), C.c_float(eps), self.adam_step, C.c_float(0.0),
                                      None, None, 1, l.b.numel(), s), "adam")

    def infer(self, pos, vel_hist, cls, scene=None, nscene: int = 1):
        """One inference pass: graph (fixed-capacity rows where N x capacity fits
        int32 offsets, else CSR), features, network; returns y [N, 3]."""
        N = pos.shape[0]
        r = self.graph_rows(pos, scene, nscene)
        fused = self.kstep == 32
        x0 = None if fused else self.features(vel_hist, cls, pos, train=False)
        feats = (vel_hist, cls) if fused else None
        # (edge sums in the graph's cell order, forward(order=...), measured no faster: node order)
        if r is not None:
            ci, deg, cap, eps = r
            return self.forward(x0, pos, None, ci, N, train=False, eps=eps, rows=(deg, cap), feats=feats)
        rp, ci, _ = self.graph(pos, scene, nscene, eps=True)
        return self.forward(x0, pos, rp, ci, N, train=False, eps=self.graph_eps, feats=feats)

    def rollout_step(self, pos, vel_hist, y, cls, body, nbodies: int, h: float, body_sums) -> None:
        """x, v history advanced by stride h from the network's output y."""
        _chk(nnlib().dmph_nn_rollout_step(_p(pos), _p(vel_hist), _p(y), C.c_float(self.ascale), _p(cls), _p(body),
                                          _p(body_sums), nbodies, C.c_float(h), pos.shape[0], self._s()), "rollout")


# ---------------------------------------------------------------------------------------------
# trajectories of the plain solver, training and rollouts
# ---------------------------------------------------------------------------------------------

STRIDE = 10   # solver steps per surrogate step


class Trajectories:
    """Frames of a batch of scenes from the plain FP32 solver (rigid debris) at
    the given solver steps: positions and velocities of every particle, original
    order, FP32, exported on the GPU (dmph_export_state); class (0 fluid, 1
    debris, 2 wall), global body ids and the scene of each particle.
    `samples` lists the frames usable as a training sample (the 4 frames before
    it and the one after it are `stride` steps apart)."""

    def __init__(self, scenes, params, steps: List[int], stride: int = STRIDE, device: str = "cuda:0",
                 precision: Optional[int] = None):
        import torch
        import dem_mph_amd as D
        self.torch = torch
        dev = torch.device(device)
        t0 = _now()
        self.precision = D.PREC_FP32 if precision is None else precision   # FP64: the parity solver
        s = D.Solver(D.Params.from_dict(params), scenes, precision=self.precision, rigid=True)
        s.autotune(3)
        self.off = np.asarray(s.scene_offsets)
        self.dt = params["Dt"]
        self.stride = stride
        self.nscene = len(scenes)
        typ = np.concatenate([sc.type for sc in scenes])
        self.cls_np = np.where(typ < 2, 0, np.where(typ < 4, 1, 2)).astype(np.int32)
        body, boff = [], 0
        for sc in scenes:
            b = sc.body if sc.body is not None else np.full(sc.n, -1, np.int32)
            body.append(np.where(b >= 0, b + boff, -1).astype(np.int32))
            boff += int(b.max()) + 1 if (b >= 0).any() else 0
        self.body_np = np.concatenate(body)
        self.nbodies = boff
        self.scene_np = np.repeat(np.arange(len(scenes), dtype=np.int32), np.diff(self.off))
        self.steps = list(steps)
        nf = len(self.steps)
        self.pos = torch.empty((nf, s.n, 3), dtype=torch.float32, device=dev)
        self.vel = torch.empty((nf, s.n, 3), dtype=torch.float32, device=dev)
        cur = 0
        for f, st in enumerate(self.steps):
            if st > cur:
                s.step(st - cur)
                cur = st
            s.export_state(self.pos[f].data_ptr(), self.vel[f].data_ptr())
        s.sync()   # raises if a neighbour row overflowed its capacity: the frames would be invalid
        self.sim_seconds = _now() - t0
        self.solver_steps = cur
        s.close()
        self.cls = torch.tensor(self.cls_np, device=dev)
        self.body = torch.tensor(self.body_np, device=dev)
        self.scene = torch.tensor(self.scene_np, device=dev)
        self.nframes = nf
        self.samples = sample_frames(self.steps, stride)

    def centroids(self, pos_np: np.ndarray) -> np.ndarray:
        """Debris-body centroids [nbodies, 3] of positions in original order."""
        m = self.body_np >= 0
        c = np.zeros((self.nbodies, 3))
        np.add.at(c, self.body_np[m], pos_np[m])
        return c / np.bincount(self.body_np[m], minlength=self.nbodies)[:, None]


def sample_frames(steps: List[int], stride: int = STRIDE) -> List[int]:
    """Frames usable as training samples: the HIST - 1 frames before and the
    one after are `stride` solver steps apart."""
    return [f for f in range(HIST - 1, len(steps) - 1)
            if all(steps[f - t] - steps[f - t - 1] == stride for t in range(HIST - 1))
            and steps[f + 1] - steps[f] == stride]


def _now() -> float:
    import time
    return time.perf_counter()


def scales(traj: Trajectories) -> Tuple[float, float]:
    """Feature and target scales from the data: 1 / std of moving particles'
    velocities and of their 10 dt accelerations."""
    torch = traj.torch
    mv = traj.cls < 2
    v = traj.vel[:, mv]
    a = (traj.vel[1:, mv] - traj.vel[:-1, mv]) / (traj.stride * traj.dt)
    return 1.0 / max(float(v.std()), 1e-6), 1.0 / max(float(a.std()), 1e-6)


DEBRIS_WEIGHT = 5   # loss weight of debris particles (the rollout metric is the debris centroids)


def _batch(traj: Trajectories, scenes: List[int], frame: int, ascale: float, noise: float = 0.0, gen=None):
    """Inputs of `scenes` (indices into traj) at `frame`: positions, velocity
    history, class, body, scene ids (0..len-1), target, per-node loss weight.
    With `noise` (m/s), moving particles' history gets Gaussian noise and the
    target is the acceleration from the noisy newest velocity to the next
    frame's, so the network learns to pull a drifted state back."""
    torch = traj.torch
    idx = torch.cat([torch.arange(int(traj.off[k]), int(traj.off[k + 1]), device=traj.pos.device) for k in scenes])
    pos = traj.pos[frame, idx].contiguous()
    vh = torch.stack([traj.vel[frame - t, idx] for t in range(HIST)])
    cls = traj.cls[idx].contiguous()
    if noise > 0.0:
        mv = (cls < 2).to(torch.float32)[None, :, None]
        vh = vh + noise * mv * torch.randn(vh.shape, generator=gen, device=vh.device)
    vh = vh.contiguous()
    sc = torch.cat([torch.full((int(traj.off[k + 1] - traj.off[k]),), q, dtype=torch.int32, device=pos.device)
                    for q, k in enumerate(scenes)])
    tgt = ((traj.vel[frame + 1, idx] - vh[0]) * (ascale / (traj.stride * traj.dt))).contiguous()
    weight = torch.where(cls == 0, 1, torch.where(cls == 1, DEBRIS_WEIGHT, 0)).to(torch.int32)
    return pos, vh, cls, sc, tgt, weight, idx


def window_steps(start: int, gap: int, windows: int, stride: int = STRIDE) -> List[int]:
    """Solver steps of `windows` training windows of HIST + 1 frames each, `gap` steps apart."""
    return [start + w * gap + k * stride for w in range(windows) for k in range(HIST + 1)]


def train(sur: Surrogate, traj: Trajectories, iters: int, batch: int = 4, lr: float = 1e-3, seed: int = 0,
          log=None, noise: float = 0.0) -> List[float]:
    """Adam on random (scene batch, sample frame) pairs; cosine learning-rate decay."""
    rng = np.random.default_rng(seed)
    gen = traj.torch.Generator(device=traj.pos.device)
    gen.manual_seed(seed)
    losses = []
    for it in range(iters):
        sc = sorted(rng.choice(traj.nscene, size=min(batch, traj.nscene), replace=False).tolist())
        fr = int(traj.samples[int(rng.integers(len(traj.samples)))])
        pos, vh, cls, scn, tgt, wgt, _ = _batch(traj, sc, fr, sur.ascale, noise, gen)
        n = pos.shape[0]
        rp, ci, _ = sur.graph(pos, scn, len(sc))
        x0 = sur.features(vh, cls, pos)
        sur.forward(x0, pos, rp, ci, n)
        # targets in units of their standard deviation, clipped at 5 (DEM contact impulses are heavy-tailed)
        losses.append(sur.loss_and_backward(x0, pos, rp, ci, n, tgt, wgt, int(wgt.sum().item()), clip=5.0))
        sur.adam(lr * 0.5 * (1.0 + np.cos(np.pi * it / max(iters, 1))))
        if log and (it % 50 == 0 or it == iters - 1):
            log("iter %d loss %.4f" % (it, losses[-1]))
    return losses


def rollout(sur: Optional[Surrogate], traj: Trajectories, start: int, steps: int, reanchor: int, tol: float,
            baseline: Optional[str] = None):
    """Roll every scene of `traj` forward `steps` surrogate steps from `start`
    (or with a baseline without the network: "still" a = 0, "gravity" a = g).
    Every `reanchor` steps the state is reset to the plain solver's frame; a
    scene whose debris-centroid RMS error at a check exceeds `tol` falls back to
    the plain solver for the rest.  Returns per-scene errors and timings."""
    torch = traj.torch
    dev = traj.pos.device
    pos = traj.pos[start].clone()
    vh = torch.stack([traj.vel[start - t] for t in range(HIST)]).contiguous()
    n = pos.shape[0]
    body_sums = torch.zeros((max(traj.nbodies, 1), 4), dtype=torch.float32, device=dev)
    fell = np.zeros(traj.nscene, bool)
    err_debris = np.zeros(traj.nscene)
    err_fluid = np.zeros(traj.nscene)
    h = traj.stride * traj.dt
    g = np.array([0.0, 0.0, -9.81])
    t_net = 0.0
    for k in range(1, steps + 1):
        t0 = _now()
        if baseline is None:
            y = sur.infer(pos, vh, traj.cls, traj.scene, traj.nscene)
            ascale = sur.ascale
        else:
            a = np.zeros(3) if baseline == "still" else g
            y = torch.tensor(np.tile(a, (n, 1)).astype(np.float32), device=dev)
            ascale = 1.0
        _chk(nnlib().dmph_nn_rollout_step(_p(pos), _p(vh), _p(y), C.c_float(ascale), _p(traj.cls), _p(traj.body),
                                          _p(body_sums), traj.nbodies, C.c_float(h), n,
                                          C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)), "rollout")
        torch.cuda.synchronize(dev)
        t_net += _now() - t0
        if k % reanchor == 0 or k == steps:
            ref = traj.pos[start + k].cpu().numpy().astype(np.float64)
            cur = pos.cpu().numpy().astype(np.float64)
            cr, cc = traj.centroids(ref), traj.centroids(cur)
            boff = 0
            for q in range(traj.nscene):
                sl = slice(int(traj.off[q]), int(traj.off[q + 1]))
                bq = np.unique(traj.body_np[sl][traj.body_np[sl] >= 0])
                e = float(np.sqrt(((cc[bq] - cr[bq]) ** 2).sum(axis=1).mean())) if len(bq) else 0.0
                fl = traj.cls_np[sl] == 0
                ef = float(np.sqrt(((cur[sl][fl] - ref[sl][fl]) ** 2).sum(axis=1).mean())) if fl.any() else 0.0
                if not fell[q]:
                    err_debris[q] = max(err_debris[q], e)
                    err_fluid[q] = max(err_fluid[q], ef)
                    if e > tol:
                        fell[q] = True
            if k < steps:   # re-anchor on the plain solver's state
                pos.copy_(traj.pos[start + k])
                vh.copy_(torch.stack([traj.vel[start + k - t] for t in range(HIST)]))
    return {"debris_rms_max": err_debris, "fluid_rms_max": err_fluid, "fell_back": fell, "seconds": t_net,
            "steps": steps}



def advance_batch(sur: Surrogate, scenes, steps: int, dt: float) -> Dict[str, float]:
    """Throughput: the surrogate alone advances a batch of scenes from their
    initial state (velocity history = initial velocities) `steps` strides of
    STRIDE * dt; returns the wall time per surrogate step and the node count."""
    import torch
    dev = sur.dev
    pos = torch.tensor(np.concatenate([sc.pos for sc in scenes]).astype(np.float32), device=dev)
    v0 = torch.tensor(np.concatenate([sc.vel for sc in scenes]).astype(np.float32), device=dev)
    vh = v0.unsqueeze(0).repeat(HIST, 1, 1).contiguous()
    typ = np.concatenate([sc.type for sc in scenes])
    cls = torch.tensor(np.where(typ < 2, 0, np.where(typ < 4, 1, 2)).astype(np.int32), device=dev)
    body, boff = [], 0
    for sc in scenes:
        b = sc.body if sc.body is not None else np.full(sc.n, -1, np.int32)
        body.append(np.where(b >= 0, b + boff, -1).astype(np.int32))
        boff += int(b.max()) + 1 if (b >= 0).any() else 0
    body = torch.tensor(np.concatenate(body), device=dev)
    scene = torch.tensor(np.repeat(np.arange(len(scenes), dtype=np.int32), [sc.n for sc in scenes]), device=dev)
    body_sums = torch.zeros((max(boff, 1), 4), dtype=torch.float32, device=dev)
    n = pos.shape[0]
    h = STRIDE * dt

    def one():
        y = sur.infer(pos, vh, cls, scene, len(scenes))
        sur.rollout_step(pos, vh, y, cls, body, boff, h, body_sums)

    one()   # warm-up (workspace, graph capacity)
    torch.cuda.synchronize(dev)
    t0 = _now()
    for _ in range(steps):
        one()
    torch.cuda.synchronize(dev)
    el = _now() - t0
    ok = bool(torch.isfinite(pos).all().item())
    return {"seconds_per_step": el / steps, "nodes": n, "finite": ok}


class _Active:
    """The surrogate's batch: a subset of an ensemble's scenes as node arrays on
    the GPU (class, body ids renumbered from 0 over the subset, scene ids
    0..m-1) and the subset's particle offsets."""

    def __init__(self, torch, dev, scenes, ids):
        self.ids = list(ids)
        sub = [scenes[k] for k in self.ids]
        self.m = len(sub)
        self.off = np.cumsum([0] + [sc.n for sc in sub]).astype(np.int64)
        n = int(self.off[-1])
        typ = np.concatenate([sc.type for sc in sub]) if sub else np.zeros(0, np.int32)
        self.cls_np = np.where(typ < 2, 0, np.where(typ < 4, 1, 2)).astype(np.int32)
        body, boff = [], 0
        for sc in sub:
            b = sc.body if sc.body is not None else np.full(sc.n, -1, np.int32)
            body.append(np.where(b >= 0, b + boff, -1).astype(np.int32))
            boff += int(b.max()) + 1 if (b >= 0).any() else 0
        self.body_np = np.concatenate(body) if sub else np.zeros(0, np.int32)
        self.nbodies = boff
        self.cls = torch.tensor(self.cls_np, device=dev)
        self.body = torch.tensor(self.body_np, device=dev)
        self.scene = torch.tensor(np.repeat(np.arange(self.m, dtype=np.int32), np.diff(self.off)), device=dev)
        self.body_sums = torch.zeros((max(boff, 1), 4), dtype=torch.float32, device=dev)
        self.n = n

    def rows(self, k: int) -> slice:
        q = self.ids.index(k)
        return slice(int(self.off[q]), int(self.off[q + 1]))


def _with_state(sc, pos: np.ndarray, vel: np.ndarray, time: float):
    """A copy of scene `sc` (types, rigid bodies, densities, domain) at another state."""
    import copy
    out = copy.copy(sc)
    out.pos = np.ascontiguousarray(pos, np.float64)
    out.vel = np.ascontiguousarray(vel, np.float64)
    out.time = time
    return out


def debris_centroids(sc, pos: np.ndarray) -> np.ndarray:
    """Centroids [bodies, 3] of a scene's rigid debris bodies at positions `pos`."""
    if sc.body is None or not (sc.body >= 0).any():
        return np.zeros((0, 3))
    m = sc.body >= 0
    nb = int(sc.body.max()) + 1
    c = np.zeros((nb, 3))
    np.add.at(c, sc.body[m], pos[m])
    return c / np.maximum(np.bincount(sc.body[m], minlength=nb), 1)[:, None]


def _rms(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.sqrt(((a - b) ** 2).sum(axis=1).mean())) if len(a) else 0.0


def hybrid(sur: Surrogate, scenes, params, steps: int, reanchor: int = 200, probe: int = HIST - 1,
           tol: float = 0.008) -> Dict[str, object]:
    """The surrogate with its plain-solver fallback as a deployment runs it,
    from t = 0, on one GPU, for `steps` strides (each STRIDE dt) after the
    warm-up, the probes' strides included:

    1. the plain FP32 solver advances every scene (HIST - 1) strides, writing
       the velocity history the network reads;
    2. the surrogate advances the scenes it holds; every `reanchor` strides
       it is re-anchored and judged: a solver batch is built from the
       surrogate's own state (no other trajectory exists) and advanced `probe`
       strides while the surrogate, from the same state, predicts the same
       strides; the debris-centroid RMS difference after them, extrapolated
       linearly to a re-anchor period (x reanchor / probe), is the judge's
       estimate of the error the surrogate would add before the next check;
    3. a scene whose estimate exceeds `tol`, or whose state the solver cannot
       take (a neighbour row overflows: packed denser than the physics allows),
       falls back: the solver recomputes it from its last anchored solver state
       (the surrogate's strides since are discarded), it joins the fallback
       solver batch and stays there to the end; the others resume from the
       solver's state after the probe (positions and the `probe` + 1 frames of
       history), which becomes their new anchor;
    4. the fallback batch advances STRIDE solver steps per surrogate stride.

    Every part is inside the clock (set-up of the solver batches included).
    Returns the wall time, the simulated time, per-scene final states (host,
    FP64) and the judge's records."""
    import torch
    import dem_mph_amd as D
    dev = sur.dev
    dt = params["Dt"]
    h = STRIDE * dt
    plain = dict(precision=D.PREC_FP32, compensated=False, rigid_precision=D.PREC_FP32, rigid=True)
    prm = D.Params.from_dict(params)
    sync = lambda: torch.cuda.synchronize(dev)
    t_start = _now()
    act = _Active(torch, dev, scenes, range(len(scenes)))
    N = act.n
    # 1. warm-up on the solver: HIST frames STRIDE steps apart
    fp = torch.empty((HIST, N, 3), dtype=torch.float32, device=dev)
    fv = torch.empty((HIST, N, 3), dtype=torch.float32, device=dev)
    s = D.Solver(prm, scenes, **plain)
    for f in range(HIST):
        if f:
            s.step(STRIDE)
        s.export_state(fp[f].data_ptr(), fv[f].data_ptr())
    s.sync()
    s.close()
    pos = fp[HIST - 1].clone()
    vh = torch.stack([fv[HIST - 1 - t] for t in range(HIST)]).contiguous()
    done = HIST - 1                       # strides simulated
    # the last anchored (solver) state of each scene: stride and host copy
    P0, V0 = fp[HIST - 1].cpu().numpy().astype(np.float64), fv[HIST - 1].cpu().numpy().astype(np.float64)
    anchor = {k: (done, (P0[act.rows(k)], V0[act.rows(k)])) for k in act.ids}
    del fp, fv, P0, V0
    fb, fb_ids = None, []
    judged = []
    t_solver = t_sur = 0.0

    def advance_fallback(nstrides):
        nonlocal t_solver
        if fb is not None and nstrides > 0:
            t0 = _now()
            fb.step(STRIDE * nstrides)
            t_solver += _now() - t0

    def host_state(k, P, V, rows):
        return P[rows].astype(np.float64), V[rows].astype(np.float64)

    def rebuild_fallback(new_states):
        """fallback batch = its scenes at their current state + `new_states` {scene: (pos, vel)}"""
        nonlocal fb, fb_ids
        states = dict(new_states)
        if fb is not None:
            fb.sync()
            P, V = fb.field("pos"), fb.field("vel")
            for q, k in enumerate(fb_ids):
                sl = slice(int(fb.scene_offsets[q]), int(fb.scene_offsets[q + 1]))
                states[k] = (P[sl], V[sl])
            fb.close()
        fb_ids = sorted(states)
        fb = D.Solver(prm, [_with_state(scenes[k], *states[k], done * h) for k in fb_ids], **plain)

    while done < (HIST - 1) + steps:
        left = (HIST - 1) + steps - done
        n = min(reanchor, left)
        t0 = _now()
        for _ in range(n):
            if act.m:
                y = sur.infer(pos, vh, act.cls, act.scene, act.m)
                sur.rollout_step(pos, vh, y, act.cls, act.body, act.nbodies, h, act.body_sums)
            if fb is not None:
                fb.step(STRIDE)
        sync()
        t_sur += _now() - t0
        done += n
        if done >= (HIST - 1) + steps or act.m == 0:
            continue
        # 2. re-anchor and judge the surrogate's scenes
        P = pos.cpu().numpy()
        V = vh[0].cpu().numpy()
        finite = {k: bool(np.isfinite(P[act.rows(k)]).all() and np.isfinite(V[act.rows(k)]).all()) for k in act.ids}
        probe_ids = [k for k in act.ids if finite[k]]
        lost = [k for k in act.ids if not finite[k]]
        est, fall = {}, set(lost)
        frames = {}   # scene -> (positions, velocities) [probe + 1, n_k, 3] of the solver's probe
        if probe_ids:
            t0 = _now()
            p2, v2 = pos.clone(), vh.clone()
            for _ in range(probe):
                y = sur.infer(p2, v2, act.cls, act.scene, act.m)
                sur.rollout_step(p2, v2, y, act.cls, act.body, act.nbodies, h, act.body_sums)
            sync()
            P2 = p2.cpu().numpy()
            del p2, v2
            t_sur += _now() - t0
            t0 = _now()

            def run_probe(ids):
                ps = D.Solver(prm, [_with_state(scenes[k], *host_state(k, P, V, act.rows(k)), done * h) for k in ids],
                              **plain)
                try:
                    fpp = torch.empty((probe + 1, ps.n, 3), dtype=torch.float32, device=dev)
                    fvv = torch.empty((probe + 1, ps.n, 3), dtype=torch.float32, device=dev)
                    ps.export_state(fpp[0].data_ptr(), fvv[0].data_ptr())
                    for f in range(1, probe + 1):
                        ps.step(STRIDE)
                        ps.export_state(fpp[f].data_ptr(), fvv[f].data_ptr())
                    ps.sync()   # raises when a row overflowed: a state too dense to be physical
                    off_ = ps.scene_offsets
                    return {k: (fpp[:, int(off_[q]):int(off_[q + 1])], fvv[:, int(off_[q]):int(off_[q + 1])])
                            for q, k in enumerate(ids)}
                finally:
                    ps.close()

            try:
                frames = run_probe(probe_ids)
            except D.DmphError:
                for k in probe_ids:   # which scenes: one at a time
                    try:
                        frames.update(run_probe([k]))
                    except D.DmphError:
                        lost.append(k)
                        fall.add(k)
            t_solver += _now() - t0
            for k, (fpp, _) in frames.items():
                c_sol = debris_centroids(scenes[k], fpp[probe].cpu().numpy().astype(np.float64))
                c_sur = debris_centroids(scenes[k], P2[act.rows(k)].astype(np.float64))
                est[k] = _rms(c_sur, c_sol) * reanchor / probe
                if est[k] > tol:
                    fall.add(k)
        judged.append({"stride": done, "estimate_m": {int(k): v for k, v in est.items()},
                       "centroids": {int(k): debris_centroids(scenes[k], P[act.rows(k)].astype(np.float64)).tolist()
                                     for k in frames},
                       "fell_back": sorted(int(k) for k in fall), "invalid_state": sorted(int(k) for k in lost)})
        advance_fallback(probe)   # the fallback batch keeps pace with the probe
        done += probe
        t0 = _now()
        if fall:
            # 3. the solver takes over: each fallen scene is recomputed from its last
            # anchored solver state (the surrogate's strides since are discarded);
            # every scene on the surrogate was anchored at the same stride
            a_done = {anchor[k][0] for k in fall}
            assert len(a_done) == 1
            a_done = a_done.pop()
            ids = sorted(fall)
            cs = D.Solver(prm, [_with_state(scenes[k], *anchor[k][1], a_done * h) for k in ids], **plain)
            cs.step(STRIDE * (done - a_done))
            cs.sync()
            Pc, Vc = cs.field("pos"), cs.field("vel")
            newstates = {k: (Pc[int(cs.scene_offsets[q]):int(cs.scene_offsets[q + 1])],
                             Vc[int(cs.scene_offsets[q]):int(cs.scene_offsets[q + 1])]) for q, k in enumerate(ids)}
            cs.close()
            rebuild_fallback(newstates)
        # the surrogate resumes from the solver's state for the scenes it keeps
        keep = [k for k in act.ids if k in frames and k not in fall]
        if keep:
            pos = torch.cat([frames[k][0][probe] for k in keep]).contiguous()
            vh = torch.stack([torch.cat([frames[k][1][probe - t] for k in keep]) for t in range(HIST)]).contiguous()
            for k in keep:
                anchor[k] = (done, (frames[k][0][probe].cpu().numpy().astype(np.float64),
                                    frames[k][1][probe].cpu().numpy().astype(np.float64)))
        act = _Active(torch, dev, scenes, keep)
        if not keep:
            pos = torch.empty((0, 3), dtype=torch.float32, device=dev)
        t_solver += _now() - t0
    sync()
    if fb is not None:
        fb.sync()
    elapsed = _now() - t_start
    # final states, host FP64, per scene
    final = {}
    if act.m:
        P = pos.cpu().numpy().astype(np.float64)
        for k in act.ids:
            final[k] = P[act.rows(k)]
    if fb is not None:
        P = fb.field("pos")
        for q, k in enumerate(fb_ids):
            final[k] = P[int(fb.scene_offsets[q]):int(fb.scene_offsets[q + 1])]
        fb.close()
    return {"seconds": elapsed, "surrogate_seconds": t_sur, "solver_seconds": t_solver,
            "strides": done, "solver_steps_equivalent": done * STRIDE, "simulated_s": done * h,
            "on_surrogate_at_end": len(act.ids), "fell_back": len(fb_ids), "judged": judged, "final": final}


def hybrid_report(sur: Surrogate, scenes, params, steps: int, reanchor: int = 200, tol: float = 0.008,
                  reference_precision: Optional[int] = None) -> Dict[str, object]:
    """hybrid() timed against the plain FP32 solver batch advancing the same
    scenes over the same simulated time from t = 0 (autotuned, outside the
    clock), and its accuracy against a reference solver trajectory (the
    mixed-precision solver unless `reference_precision` is FP64): debris-
    centroid RMS per scene at the end and, at each judgement, the true error of
    the surrogate's state next to the judge's estimate."""
    import torch
    import dem_mph_amd as D
    r = hybrid(sur, scenes, params, steps, reanchor=reanchor, tol=tol)
    total = r["solver_steps_equivalent"]
    dt = params["Dt"]
    prm = D.Params.from_dict(params)
    s = D.Solver(prm, scenes, precision=D.PREC_FP32, compensated=False, rigid_precision=D.PREC_FP32, rigid=True)
    s.autotune(2)   # on the initial state, outside the clock
    s.sync()
    t0 = _now()
    s.step(total)
    s.sync()
    t_plain = _now() - t0
    s.close()
    # reference: frames at every judgement and at the end
    marks = [j["stride"] * STRIDE for j in r["judged"]] + [total]
    prec = D.PREC_FP32 if reference_precision is None else reference_precision
    ref = Trajectories(scenes, params, marks, precision=prec)
    off = ref.off
    Pend = ref.pos[len(marks) - 1].cpu().numpy().astype(np.float64)
    err = np.zeros(len(scenes))
    for k, sc in enumerate(scenes):
        err[k] = _rms(debris_centroids(sc, r["final"][k]), debris_centroids(sc, Pend[int(off[k]):int(off[k + 1])]))
    judge = []
    for f, j in enumerate(r["judged"]):
        Pf = ref.pos[f].cpu().numpy().astype(np.float64)
        for k, c in j["centroids"].items():
            true = _rms(np.asarray(c), debris_centroids(scenes[k], Pf[int(off[k]):int(off[k + 1])]))
            judge.append({"stride": j["stride"], "scene": k, "true_error_m": true, "estimate_m": j["estimate_m"][k],
                          "fell_back": k in j["fell_back"]})
    del ref
    sim = r["simulated_s"]
    fb_set = set(k for j in r["judged"] for k in j["fell_back"])
    return {"scenes": len(scenes), "particles": int(sum(sc.n for sc in scenes)), "surrogate_strides": steps,
            "stride_dt": STRIDE, "reanchor_strides": reanchor, "probe_strides": HIST - 1, "tolerance_m": tol,
            "simulated_s": sim, "seconds": r["seconds"], "surrogate_seconds": r["surrogate_seconds"],
            "solver_seconds": r["solver_seconds"], "sim_s_per_wall_s": sim / r["seconds"],
            "plain_batched_seconds": t_plain, "plain_batched_sim_s_per_wall_s": sim / t_plain,
            "vs_plain_batched": t_plain / r["seconds"],
            "fell_back": r["fell_back"], "on_surrogate_at_end": r["on_surrogate_at_end"],
            "invalid_states": sum(len(j["invalid_state"]) for j in r["judged"]),
            "reference": "mixed-precision solver" if prec == D.PREC_FP32 else "FP64 solver",
            "debris_centroid_rms_m_mean": float(err.mean()), "debris_centroid_rms_m_max": float(err.max()),
            "within_tolerance": int((err <= tol).sum()),
            "within_tolerance_of_surrogate_kept": int(sum(1 for k in range(len(scenes))
                                                         if k not in fb_set and err[k] <= tol)),
            "surrogate_kept": int(len(scenes) - len(fb_set)),
            "judgements": judge}


def save(sur: Surrogate, path: str, **meta) -> None:
    sd = sur.state_dict()
    np.savez(path, vscale=sur.vscale, ascale=sur.ascale, d=sur.d, tank_lo=sur.tank_lo, tank_hi=sur.tank_hi,
             meta=np.array(repr(meta)), **{"p/" + k: v.astype(np.float32) for k, v in sd.items()})


def load(path: str, kstep: int = 32) -> Surrogate:
    z = np.load(path, allow_pickle=False)
    sur = Surrogate(d=float(z["d"]), tank_lo=z["tank_lo"], tank_hi=z["tank_hi"], vscale=float(z["vscale"]),
                    ascale=float(z["ascale"]), kstep=kstep)
    sur.load_state({k[2:]: z[k].astype(np.float64) for k in z.files if k.startswith("p/")})
    return sur


def frame_errors(ref: "Trajectories", other: "Trajectories", frames: List[int]) -> Dict[str, float]:
    """Per-scene debris-centroid and fluid RMS of `other` against `ref` (same
    scenes and solver steps), the largest over `frames`, then mean / max over
    the scenes."""
    ed = np.zeros(ref.nscene)
    ef = np.zeros(ref.nscene)
    for f in frames:
        a = ref.pos[f].cpu().numpy().astype(np.float64)
        b = other.pos[f].cpu().numpy().astype(np.float64)
        ca, cb = ref.centroids(a), ref.centroids(b)
        for q in range(ref.nscene):
            sl = slice(int(ref.off[q]), int(ref.off[q + 1]))
            bq = np.unique(ref.body_np[sl][ref.body_np[sl] >= 0])
            if len(bq):
                ed[q] = max(ed[q], float(np.sqrt(((cb[bq] - ca[bq]) ** 2).sum(axis=1).mean())))
            fl = ref.cls_np[sl] == 0
            if fl.any():
                ef[q] = max(ef[q], float(np.sqrt(((b[sl][fl] - a[sl][fl]) ** 2).sum(axis=1).mean())))
    return {"debris_centroid_rms_m_mean": float(ed.mean()), "debris_centroid_rms_m_max": float(ed.max()),
            "fluid_rms_m_mean": float(ef.mean())}


def evaluate(sur: Surrogate, scenes, params, start: int, steps: int, reanchor: int, tol: float,
             fp64: bool = False) -> Dict[str, object]:
    """Held-out scenes: reference frames from the mixed-precision solver (or,
    with `fp64`, from the FP64 parity solver, and then also the mixed solver's
    own error against it over the same frames), then the surrogate's rollout and
    the two baselines without a network."""
    import dem_mph_amd as D
    frames = [start - (HIST - 1) * STRIDE + k * STRIDE for k in range(steps + HIST)]
    traj = Trajectories(scenes, params, frames, precision=D.PREC_FP64 if fp64 else None)
    out = {"scenes": len(scenes), "particles": int(traj.off[-1]), "start_time_s": start * traj.dt,
           "horizon_s": steps * STRIDE * traj.dt, "reanchor_steps": reanchor, "tolerance_m": tol,
           "reference": "FP64 solver" if fp64 else "mixed-precision solver"}
    if fp64:
        mixed = Trajectories(scenes, params, frames)
        out["mixed_solver"] = frame_errors(traj, mixed, list(range(HIST - 1, len(frames))))
        del mixed
    for name, base in (("surrogate", None), ("still", "still"), ("gravity", "gravity")):
        r = rollout(sur if base is None else None, traj, HIST - 1, steps, reanchor, tol, baseline=base)
        out[name] = {"debris_centroid_rms_m_mean": float(r["debris_rms_max"].mean()),
                     "debris_centroid_rms_m_max": float(r["debris_rms_max"].max()),
                     "fluid_rms_m_mean": float(r["fluid_rms_max"].mean()),
                     "fell_back": int(r["fell_back"].sum()), "seconds": r["seconds"]}
    return out


def main(argv=None) -> int:
    import argparse
    import json
    import sys
    from . import scenes as S
    ap = argparse.ArgumentParser(description="train and evaluate the surrogate on C3-style scenes")
    ap.add_argument("--train-scenes", type=int, default=128)
    ap.add_argument("--eval-scenes", type=int, default=32)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--windows", type=int, default=16)
    ap.add_argument("--gap", type=int, default=300, help="solver steps between training windows")
    ap.add_argument("--iters", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--noise", type=float, default=0.01, help="velocity-history noise in training (m/s)")
    ap.add_argument("--eval-start", type=int, default=2000, help="solver step where rollouts start")
    ap.add_argument("--eval-steps", type=int, default=100, help="surrogate steps (10 solver steps each)")
    ap.add_argument("--reanchor", type=int, default=200)
    ap.add_argument("--tol", type=float, default=0.008, help="debris-centroid RMS tolerance (1%% of the tank length)")
    ap.add_argument("--out", default=os.path.join(PKG, "surrogate_c3.npz"))
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--throughput", type=int, default=0, metavar="SCENES",
                    help="only time the saved model's rollout on SCENES ensemble scenes (seed 1), 3 steps")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first one device-to-device copy of MIB MiB (counter calibration)")
    a = ap.parse_args(argv)
    log = lambda m: print("surrogate: " + m, file=sys.stderr, flush=True)
    params = S.params_c3()
    if a.copy_probe:
        from .profiling import copy_probe
        copy_probe(a.copy_probe)
    if a.throughput:
        sc = S.ensemble_c3(a.throughput, seed=1)
        r = advance_batch(load(a.out), sc, 3, params["Dt"])
        print(json.dumps(r))
        return 0
    rep = {}
    if not a.eval_only:
        sc = S.ensemble_c3(a.train_scenes, seed=a.seed)
        log("training data: %d scenes, %d windows" % (len(sc), a.windows))
        traj = Trajectories(sc, params, window_steps(HIST * STRIDE, a.gap, a.windows))
        vs, asc = scales(traj)
        log("data %.1f s; vscale %.3g ascale %.3g" % (traj.sim_seconds, vs, asc))
        sur = Surrogate(d=0.01, tank_lo=[0, 0, 0], tank_hi=[0.8, 0.4, 0.6], vscale=vs, ascale=asc, seed=a.seed)
        t0 = _now()
        losses = train(sur, traj, a.iters, a.batch, a.lr, a.seed, log, noise=a.noise)
        rep["train"] = {"scenes": len(sc), "windows": a.windows, "iters": a.iters, "batch": a.batch, "noise": a.noise,
                        "seconds": _now() - t0, "loss_first": losses[0], "loss_last10": float(np.mean(losses[-10:]))}
        save(sur, a.out, **rep["train"])
        del traj
    else:
        sur = load(a.out)
    ev = S.ensemble_c3(a.eval_scenes, seed=a.seed + 1000)   # held-out seeds
    rep["eval"] = evaluate(sur, ev, params, a.eval_start, a.eval_steps, a.reanchor, a.tol)
    print(json.dumps(rep))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
