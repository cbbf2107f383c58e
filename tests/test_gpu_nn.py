"""Surrogate kernels on the GPU (libdemmph_nn.so through its C ABI) against
numpy and the FP64 network oracle (oracle/nn_oracle.py)."""
import ctypes as C

import numpy as np
import pytest

import nn_oracle as NO

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _bf16_bits(a: np.ndarray) -> np.ndarray:
    """round-to-nearest-even bf16 bits of float32 values"""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def _bf16_float(bits: np.ndarray) -> np.ndarray:
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _dev_bf16(a):
    return torch.tensor(_bf16_bits(a).view(np.int16), device="cuda:0")


def _host_bf16(t):
    return _bf16_float(t.cpu().numpy().view(np.uint16))


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


@pytest.fixture(scope="module")
def S():
    from dem_mph_amd import surrogate as S
    return S


@pytest.mark.parametrize("kstep", [16, 32])
@pytest.mark.parametrize("K,N,relu,res,half,M", [(32, 128, 1, 0, 0, 300), (128, 128, 0, 1, 0, 300),
                                                 (256, 128, 1, 0, 0, 300), (128, 256, 0, 0, 0, 300),
                                                 (128, 256, 0, 0, 1, 300), (256, 128, 1, 1, 1, 300),
                                                 (128, 128, 0, 1, 0, 200_001), (256, 128, 1, 0, 0, 200_001)])
def test_gemm_matches_numpy(S, kstep, K, N, relu, res, half, M):
    """bf16 (or IEEE half, flag DMPH_NN_F16_OUT) output against numpy on the
    same bf16 inputs; M = 300 leaves a partial row tile, M = 200,001 makes the
    persistent workgroups walk several tiles each, the last one partial."""
    rng = np.random.default_rng(K + N + kstep)
    X = rng.standard_normal((M, K)).astype(np.float32)
    Wt = (rng.standard_normal((N, K)) / np.sqrt(K)).astype(np.float32)
    b = rng.standard_normal(N).astype(np.float32)
    R = rng.standard_normal((M, N)).astype(np.float32)
    Xd, Wd, Rd = _dev_bf16(X), _dev_bf16(Wt), _dev_bf16(R)
    bd = torch.tensor(b, device="cuda:0")
    Y = torch.zeros((M, N), dtype=torch.int16, device="cuda:0")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    flags = (S.GEMM_RELU if relu else 0) | (S.GEMM_F16_OUT if half else 0)
    rc = S.nnlib().dmph_nn_gemm(_p(Xd), K, _p(Wd), _p(bd), _p(Rd) if res else None, N, _p(Y), N, M, K, N, flags,
                                kstep, s)
    assert rc == 0
    ref = _host_bf16(Xd) @ _host_bf16(Wd).T + b + (_host_bf16(Rd) if res else 0.0)
    if relu:
        ref = np.maximum(ref, 0.0)
    got = Y.cpu().numpy().view(np.float16).astype(np.float64) if half else _host_bf16(Y)
    err = np.abs(got - ref)
    # rounding of the output: 2^-9 relative (bf16), 2^-12 (half); FP32 sums in a different order
    tol = 2e-3 if half else 1e-2
    assert err.max() <= tol * np.abs(ref).max(), err.max()
    assert np.abs(got - ref).mean() <= tol / 5 * np.abs(ref).mean() + 1e-6


@pytest.mark.parametrize("kstep", [16, 32])
@pytest.mark.parametrize("M", [301, 200_003])
def test_gemm_head_matches_gemm_then_head(S, kstep, M):
    """Inference decoder tail (hidden layer + 3-output head in one kernel)
    against the two kernels the training path runs, on the same bf16 inputs;
    partial last tiles; at M = 200,003 each persistent workgroup walks several
    tiles; X with a row stride of 256."""
    rng = np.random.default_rng(kstep)
    X = rng.standard_normal((M, 256)).astype(np.float32)
    Wt = (rng.standard_normal((128, 128)) / np.sqrt(128)).astype(np.float32)
    b = rng.standard_normal(128).astype(np.float32)
    Wd = rng.standard_normal((3, 128)).astype(np.float32)
    bd = rng.standard_normal(3).astype(np.float32)
    Xd, Wdv = _dev_bf16(X), _dev_bf16(Wt)
    bdev, Wdd, bdd = (torch.tensor(a, device="cuda:0") for a in (b, Wd, bd))
    lib = S.nnlib()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    G = torch.zeros((M, 128), dtype=torch.int16, device="cuda:0")
    y1 = torch.zeros((M, 3), dtype=torch.float32, device="cuda:0")
    y2 = torch.full((M, 3), float("nan"), dtype=torch.float32, device="cuda:0")
    assert lib.dmph_nn_gemm(_p(Xd), 256, _p(Wdv), _p(bdev), None, 0, _p(G), 128, M, 128, 128, S.GEMM_RELU, kstep, s) == 0
    assert lib.dmph_nn_head(_p(G), _p(Wdd), _p(bdd), _p(y1), M, s) == 0
    assert lib.dmph_nn_gemm_head(_p(Xd), 256, _p(Wdv), _p(bdev), _p(Wdd), _p(bdd), _p(y2), M, kstep, s) == 0
    a, c = y1.cpu().numpy(), y2.cpu().numpy()
    assert np.isfinite(c).all()
    np.testing.assert_allclose(c, a, rtol=1e-5, atol=1e-5 * np.abs(a).max())


@pytest.mark.parametrize("M", [301, 200_003])
def test_node_update_matches_three_gemms(S, M):
    """dmph_nn_node (u, h' and the next [A | B] in registers, W2 and Wab staged
    with permuted columns) against dmph_nn_gemm run three times on the same
    bf16 inputs: h' and AB agree to the rounding of their stored values."""
    rng = np.random.default_rng(M)
    HS = rng.standard_normal((M, 256)).astype(np.float32)
    W1 = (rng.standard_normal((128, 256)) / 16).astype(np.float32)
    W2 = (rng.standard_normal((128, 128)) / 11).astype(np.float32)
    Wab = (rng.standard_normal((256, 128)) / 11).astype(np.float32)
    b1, b2 = (torch.tensor(rng.standard_normal(128).astype(np.float32), device="cuda:0") for _ in range(2))
    HSd, W1d, W2d, Wabd = _dev_bf16(HS), _dev_bf16(W1), _dev_bf16(W2), _dev_bf16(Wab)
    lib = S.nnlib()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    u = torch.zeros((M, 128), dtype=torch.int16, device="cuda:0")
    h1 = torch.zeros((M, 128), dtype=torch.int16, device="cuda:0")
    ab1 = torch.zeros((M, 256), dtype=torch.int16, device="cuda:0")
    assert lib.dmph_nn_gemm(_p(HSd), 256, _p(W1d), _p(b1), None, 0, _p(u), 128, M, 256, 128, S.GEMM_RELU, 32, s) == 0
    assert lib.dmph_nn_gemm(_p(u), 128, _p(W2d), _p(b2), _p(HSd), 256, _p(h1), 128, M, 128, 128, 0, 32, s) == 0
    assert lib.dmph_nn_gemm(_p(h1), 128, _p(Wabd), None, None, 0, _p(ab1), 256, M, 128, 256, S.GEMM_F16_OUT, 32, s) == 0
    h2 = torch.zeros((M, 256), dtype=torch.int16, device="cuda:0")   # h' into the left half of a [h | S] row
    ab2 = torch.full((M, 256), -1, dtype=torch.int16, device="cuda:0")
    assert lib.dmph_nn_node(_p(HSd), 256, _p(W1d), _p(b1), _p(W2d), _p(b2), _p(Wabd), _p(h2), 256, _p(ab2), M, None,
                            s) == 0
    a, b = _host_bf16(h1), _host_bf16(h2[:, :128])
    assert not h2[:, 128:].any()
    assert np.abs(a - b).max() <= 2e-2 * np.abs(a).max() and np.abs(a - b).mean() <= 1e-3 * np.abs(a).mean()
    x = ab1.cpu().numpy().view(np.float16).astype(np.float64)
    y = ab2.cpu().numpy().view(np.float16).astype(np.float64)
    assert np.isfinite(y).all()
    assert np.abs(x - y).max() <= 2e-2 * np.abs(x).max() and np.abs(x - y).mean() <= 2e-3 * np.abs(x).mean()


@pytest.mark.parametrize("M", [301, 200_003])
def test_node_head_matches_gemms_and_head(S, M):
    """dmph_nn_node_head (last round + decoder) against dmph_nn_gemm twice and
    dmph_nn_gemm_head on the same bf16 inputs."""
    rng = np.random.default_rng(M + 1)
    HS = rng.standard_normal((M, 256)).astype(np.float32)
    W1 = (rng.standard_normal((128, 256)) / 16).astype(np.float32)
    W2 = (rng.standard_normal((128, 128)) / 11).astype(np.float32)
    Wd1 = (rng.standard_normal((128, 128)) / 11).astype(np.float32)
    HSd, W1d, W2d, Wd1d = _dev_bf16(HS), _dev_bf16(W1), _dev_bf16(W2), _dev_bf16(Wd1)
    b1, b2, bd1 = (torch.tensor(rng.standard_normal(128).astype(np.float32), device="cuda:0") for _ in range(3))
    Wd2 = torch.tensor(rng.standard_normal((3, 128)).astype(np.float32), device="cuda:0")
    bd2 = torch.tensor(rng.standard_normal(3).astype(np.float32), device="cuda:0")
    lib = S.nnlib()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    u = torch.zeros((M, 128), dtype=torch.int16, device="cuda:0")
    h = torch.zeros((M, 128), dtype=torch.int16, device="cuda:0")
    y1 = torch.zeros((M, 3), dtype=torch.float32, device="cuda:0")
    y2 = torch.full((M, 3), float("nan"), dtype=torch.float32, device="cuda:0")
    assert lib.dmph_nn_gemm(_p(HSd), 256, _p(W1d), _p(b1), None, 0, _p(u), 128, M, 256, 128, S.GEMM_RELU, 32, s) == 0
    assert lib.dmph_nn_gemm(_p(u), 128, _p(W2d), _p(b2), _p(HSd), 256, _p(h), 128, M, 128, 128, 0, 32, s) == 0
    assert lib.dmph_nn_gemm_head(_p(h), 128, _p(Wd1d), _p(bd1), _p(Wd2), _p(bd2), _p(y1), M, 32, s) == 0
    assert lib.dmph_nn_node_head(_p(HSd), 256, _p(W1d), _p(b1), _p(W2d), _p(b2), _p(Wd1d), _p(bd1), _p(Wd2), _p(bd2),
                                 _p(y2), M, s) == 0
    a, b = y1.cpu().numpy(), y2.cpu().numpy()
    assert np.isfinite(b).all()
    err = np.abs(a - b)
    assert err.max() <= 2e-2 * np.abs(a).max() and err.mean() <= 2e-3 * np.abs(a).mean()


def test_encode_matches_features_and_gemms(S):
    """dmph_nn_encode (features, both encoder layers and the first [A | B] in
    one kernel) against dmph_nn_features and dmph_nn_gemm three times, on a
    scene of 3 classes with wall distances inside and outside the clip."""
    sur, sc, pos, vh, cls, tgt, mask = _small_problem(S, seed=5)
    n = sc.n
    lib = S.nnlib()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    x0 = sur.features(vh, cls, pos, train=False)
    h1 = torch.zeros((n, 128), dtype=torch.int16, device="cuda:0")
    h0 = torch.zeros((n, 256), dtype=torch.int16, device="cuda:0")
    ab1 = torch.zeros((n, 256), dtype=torch.int16, device="cuda:0")
    L0 = sur.rounds[0]["ab"]
    assert lib.dmph_nn_gemm(_p(x0), 32, _p(sur.enc1.wb), _p(sur.enc1.b), None, 0, _p(h1), 128, n, 32, 128, S.GEMM_RELU,
                            32, s) == 0
    assert lib.dmph_nn_gemm(_p(h1), 128, _p(sur.enc2.wb), _p(sur.enc2.b), None, 0, _p(h0), 256, n, 128, 128, 0, 32, s) == 0
    assert lib.dmph_nn_gemm(_p(h0), 256, _p(L0.wb), None, None, 0, _p(ab1), 256, n, 128, 256, S.GEMM_F16_OUT, 32, s) == 0
    h0b = torch.zeros((n, 256), dtype=torch.int16, device="cuda:0")
    ab2 = torch.zeros((n, 256), dtype=torch.int16, device="cuda:0")
    assert lib.dmph_nn_encode(_p(vh), n, _p(cls), _p(pos), C.byref(sur._feat()), _p(sur.enc1.wb), _p(sur.enc1.b),
                              _p(sur.enc2.wb), _p(sur.enc2.b), _p(L0.wb), _p(h0b), 256, _p(ab2), None, s) == 0
    a, b = _host_bf16(h0[:, :128]), _host_bf16(h0b[:, :128])
    assert np.abs(a - b).max() <= 2e-2 * np.abs(a).max() and np.abs(a - b).mean() <= 1e-3 * np.abs(a).mean()
    x = ab1.cpu().numpy().view(np.float16).astype(np.float64)
    y = ab2.cpu().numpy().view(np.float16).astype(np.float64)
    assert np.abs(x - y).max() <= 2e-2 * np.abs(x).max() and np.abs(x - y).mean() <= 2e-3 * np.abs(x).mean()


@pytest.mark.parametrize("K,N", [(32, 128), (128, 128), (256, 128), (128, 256)])
def test_gemm_dw_matches_numpy(S, K, N):
    rng = np.random.default_rng(K * N)
    M = 1000
    dY = rng.standard_normal((M, N)).astype(np.float32)
    X = rng.standard_normal((M, K)).astype(np.float32)
    dYd, Xd = _dev_bf16(dY), _dev_bf16(X)
    dW = torch.zeros((N, K), dtype=torch.float32, device="cuda:0")
    db = torch.zeros(N, dtype=torch.float32, device="cuda:0")
    lib = S.nnlib()
    scr = torch.zeros(lib.dmph_nn_dw_scratch(M, K, N, 7) // 4 + 1, dtype=torch.float32, device="cuda:0")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.dmph_nn_gemm_dw(_p(dYd), N, _p(Xd), K, _p(dW), _p(db), M, K, N, 7, _p(scr), 0, s) == 0
    ref = _host_bf16(dYd).T @ _host_bf16(Xd)
    np.testing.assert_allclose(dW.cpu().numpy(), ref, rtol=1e-4, atol=1e-3)
    np.testing.assert_allclose(db.cpu().numpy(), _host_bf16(dYd).sum(axis=0), rtol=1e-4, atol=1e-3)


def test_radius_graph_matches_brute_force(S):
    rng = np.random.default_rng(5)
    pos = rng.uniform(0.0, 0.3, (3000, 3)).astype(np.float32)
    sur = S.Surrogate(d=0.01, tank_lo=[0, 0, 0], tank_hi=[0.3, 0.3, 0.3], vscale=1.0, ascale=1.0)
    posd = torch.tensor(pos, device="cuda:0")
    rp, ci, nnz = sur.graph(posd, eps=True)
    # the edge features written with the graph equal edge_eps's
    e1 = sur.graph_eps[:2 * nnz].cpu().numpy().view(np.float16).astype(np.float64)
    e2 = torch.empty(2 * ci.numel(), dtype=torch.int32, device="cuda:0")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert S.nnlib().dmph_nn_edge_eps(_p(posd), _p(rp), _p(ci), _p(e2), len(pos), C.c_float(sur.radius), s) == 0
    e2 = e2[:2 * nnz].cpu().numpy().view(np.float16).astype(np.float64)
    assert np.abs(e1 - e2).max() <= 1e-3 and np.abs(e1).max() <= 1.0 + 1e-3
    rp, ci = rp.cpu().numpy(), ci.cpu().numpy()[:nnz]
    brp, bci = NO.radius_graph(pos.astype(np.float64), sur.radius)
    # the GPU tests r^2 in FP32: pairs within a rounding of the radius may differ
    d = np.linalg.norm(pos[:, None] - pos[None], axis=-1)
    for i in range(len(pos)):
        a = set(ci[rp[i]:rp[i + 1]].tolist())
        b = set(bci[brp[i]:brp[i + 1]].tolist())
        for j in a ^ b:
            assert abs(d[i, j] - sur.radius) < 1e-6
        assert len(ci[rp[i]:rp[i + 1]]) == len(a)   # no duplicates
    assert nnz > 10 * len(pos)


def test_fixed_capacity_rows_equal_the_csr_graph(S):
    """dmph_nn_graph_rows: the rows of dmph_nn_graph, entry for entry, at
    i * row_cap; started at 16 entries the capacity grows past the longest row
    (the cut rows are reported, nothing is written past a row's end); the edge
    sums over the rows equal those over the CSR graph."""
    sur, sc, pos, vh, cls, tgt, mask = _small_problem(S, seed=4)
    n = sc.n
    rp, ci, nnz = sur.graph(pos, eps=True)
    RP, CI = rp.cpu().numpy(), ci.cpu().numpy()
    E = sur.graph_eps[:2 * nnz].cpu().numpy().reshape(-1, 2)
    ci2, deg, cap, eps2 = sur.graph_rows(pos, row_cap=16, dist=False)
    assert cap >= np.diff(RP).max() and cap % 16 == 0
    D, C2, E2 = deg[:n].cpu().numpy(), ci2.cpu().numpy().reshape(n, cap), eps2.cpu().numpy().reshape(n, cap, 2)
    np.testing.assert_array_equal(D, np.diff(RP))
    for i in range(n):
        np.testing.assert_array_equal(C2[i, :D[i]], CI[RP[i]:RP[i + 1]])
        np.testing.assert_array_equal(E2[i, :D[i]], E[RP[i]:RP[i + 1]])
    rng = np.random.default_rng(4)
    abd = torch.tensor((rng.standard_normal((n, 256)) * 0.5).astype(np.float16).view(np.int16), device="cuda:0")
    Wc = torch.tensor((rng.standard_normal((128, 4)) * 0.5).astype(np.float32), device="cuda:0")
    bc = torch.tensor((rng.standard_normal(128) * 0.2).astype(np.float32), device="cuda:0")
    lib = S.nnlib()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    o1 = torch.zeros((n, 128), dtype=torch.int16, device="cuda:0")
    o2 = torch.zeros((n, 128), dtype=torch.int16, device="cuda:0")
    assert lib.dmph_nn_edge_fwd_f16(_p(abd), _p(rp), _p(ci), _p(sur.graph_eps), _p(Wc), _p(bc), _p(o1), 128, n,
                                    None, 0, None, None, s) == 0
    assert lib.dmph_nn_edge_fwd_f16(_p(abd), None, _p(ci2), _p(eps2), _p(Wc), _p(bc), _p(o2), 128, n, _p(deg), cap,
                                    None, None, s) == 0
    assert torch.equal(o1, o2)
    # summed in the graph's cell order: the same sums
    order = sur._graph_ws["order"][:n]
    assert np.array_equal(np.sort(order.cpu().numpy()), np.arange(n))
    o3 = torch.zeros((n, 128), dtype=torch.int16, device="cuda:0")
    assert lib.dmph_nn_edge_fwd_f16(_p(abd), None, _p(ci2), _p(eps2), _p(Wc), _p(bc), _p(o3), 128, n, _p(deg), cap,
                                    _p(order), None, s) == 0
    assert torch.equal(o1, o3)
    # distances alone (for the position fold): the distance halves of eps, same rows
    ci3, deg3, cap3, d3 = sur.graph_rows(pos, row_cap=16, dist=True)
    assert sur.rows_dist and cap3 == cap and torch.equal(deg3[:n], deg[:n])
    D3 = d3.cpu().numpy().view(np.uint16)[:n * cap].reshape(n, cap)
    C3 = ci3.cpu().numpy().reshape(n, cap)
    for i in range(n):
        np.testing.assert_array_equal(C3[i, :D[i]], CI[RP[i]:RP[i + 1]])
        np.testing.assert_array_equal(D3[i, :D[i]], (E[RP[i]:RP[i + 1], 1].astype(np.uint32) >> 16).astype(np.uint16))


def test_radius_graph_refuses_more_than_int32_edges(S):
    """50,000 nodes inside one radius: 2.5e9 edges do not fit int32 row
    offsets; the build stops before writing any row and says so."""
    rng = np.random.default_rng(9)
    pos = torch.tensor(rng.uniform(0.0, 0.01, (50_000, 3)).astype(np.float32), device="cuda:0")
    sur = S.Surrogate(d=1.0, tank_lo=[0, 0, 0], tank_hi=[1, 1, 1], vscale=1.0, ascale=1.0)
    with pytest.raises(RuntimeError, match="2\\^31"):
        sur.graph(pos)
    # the same nodes at a small radius still build
    rp, ci, nnz = sur.at_spacing(0.0005).graph(pos)
    assert 0 < nnz == int(rp[-1].item())


def _small_problem(S, seed=0):
    from dem_mph_amd import scenes as SC
    sc = SC.ensemble_c3(1, seed=seed, d=0.02, interior=(0.4, 0.2, 0.3), fluid_range=(200.0, 600.0))[0]
    sur = S.Surrogate(d=sc.d, tank_lo=[0, 0, 0], tank_hi=[0.4, 0.2, 0.3], vscale=3.0, ascale=0.1, seed=seed)
    n = sc.n
    rng = np.random.default_rng(seed)
    pos = torch.tensor(sc.pos.astype(np.float32), device="cuda:0")
    vh = torch.tensor(rng.standard_normal((5, n, 3)).astype(np.float32) * 0.2, device="cuda:0")
    cls = torch.tensor(np.where(sc.type < 2, 0, np.where(sc.type < 4, 1, 2)).astype(np.int32), device="cuda:0")
    tgt = torch.tensor(rng.standard_normal((n, 3)).astype(np.float32), device="cuda:0")
    # the training weights: fluid 1, debris 5, walls 0
    mask = torch.where(cls == 0, 1, torch.where(cls == 1, 5, 0)).to(torch.int32)
    return sur, sc, pos, vh, cls, tgt, mask


def test_network_gradients_match_fp64_oracle(S):
    """bf16 MFMA forward/backward against the FP64 oracle on the same parameters,
    inputs and graph: outputs and every parameter's gradient within 5% (norm),
    gradient directions within 0.99 cosine."""
    sur, sc, pos, vh, cls, tgt, mask = _small_problem(S)
    n = sc.n
    rp, ci, nnz = sur.graph(pos)
    x0 = sur.features(vh, cls, pos)
    y = sur.forward(x0, pos, rp, ci, n).clone()
    count = int(mask.sum().item())
    lg = sur.loss_and_backward(x0, pos, rp, ci, n, tgt, mask, count)
    P = sur.state_dict()
    X0 = _host_bf16(x0[:n])
    yr, cache = NO.forward(P, X0, pos.cpu().numpy().astype(np.float64), rp.cpu().numpy(), ci.cpu().numpy()[:nnz],
                           sur.radius)
    m = mask.cpu().numpy().astype(float)
    Gr = NO.backward(P, cache, yr, tgt.cpu().numpy().astype(np.float64), m, count)
    ey = np.linalg.norm(y.cpu().numpy() - yr) / np.linalg.norm(yr)
    assert ey < 0.05, ey
    lr_ = NO.loss(yr, tgt.cpu().numpy().astype(np.float64), m, count)
    assert abs(lg - lr_) < 0.05 * lr_, (lg, lr_)
    G = sur.grads()
    worst = {}
    for k in Gr:
        a, b = G[k].ravel(), Gr[k].ravel()
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
        worst[k] = (rel, cos)
    bad = {k: v for k, v in worst.items() if not (v[0] < 0.05 and v[1] > 0.99)}
    assert not bad, bad


def test_half_edge_sums_match_fp64(S):
    """Inference edge sums (edge_eps + edge_fwd_f16: packed IEEE-half arithmetic,
    FP32 sums every 16 edges) against FP64 sums on the same half inputs; rows
    of every length (walls: short rows), lds = 256 writing into the right half
    of a row."""
    sur, sc, pos, vh, cls, tgt, mask = _small_problem(S, seed=3)
    n = sc.n
    rp0, ci0, nnz0 = sur.graph(pos)
    # rows cut to i % 71 edges (0 .. 70: empty, tail-only, batches + tail)
    RP0, CI0 = rp0.cpu().numpy(), ci0.cpu().numpy()
    lens = np.minimum(np.diff(RP0), np.arange(n) % 71)
    RP = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    CI = np.concatenate([CI0[RP0[i]:RP0[i] + lens[i]] for i in range(n)]).astype(np.int32)
    nnz = int(RP[-1])
    rp, ci = torch.tensor(RP, device="cuda:0"), torch.tensor(CI, device="cuda:0")
    rng = np.random.default_rng(3)
    ab = (rng.standard_normal((n, 256)) * 0.7).astype(np.float16)
    Wc = (rng.standard_normal((128, 4)) * 0.8).astype(np.float32)
    bc = (rng.standard_normal(128) * 0.3).astype(np.float32)
    lib = S.nnlib()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    abd = torch.tensor(ab.view(np.int16), device="cuda:0")
    Wcd, bcd = torch.tensor(Wc, device="cuda:0"), torch.tensor(bc, device="cuda:0")
    eps = torch.empty(2 * ci.numel(), dtype=torch.int32, device="cuda:0")
    out = torch.zeros((n, 256), dtype=torch.int16, device="cuda:0")
    assert lib.dmph_nn_edge_eps(_p(pos), _p(rp), _p(ci), _p(eps), n, C.c_float(sur.radius), s) == 0
    assert lib.dmph_nn_edge_fwd_f16(_p(abd), _p(rp), _p(ci), _p(eps), _p(Wcd), _p(bcd), _p(out[:, 128:]), 256, n,
                                    None, 0, None, None, s) == 0
    got = _bf16_float(out.cpu().numpy().view(np.uint16)[:, 128:])
    assert not out[:, :128].any()   # the left half untouched
    P = pos.cpu().numpy().astype(np.float64)
    A, B = ab[:, :128].astype(np.float64), ab[:, 128:].astype(np.float64)
    ref = np.zeros((n, 128))
    assert lens.min() == 0 and lens.max() == 70 and nnz > 0
    for i in range(n):
        J = CI[RP[i]:RP[i + 1]]
        d = P[J] - P[i]
        e = np.concatenate([d, np.linalg.norm(d, axis=1, keepdims=True)], axis=1) / sur.radius
        ref[i] = np.maximum(A[i] + bc + B[J] + e @ Wc.T.astype(np.float64), 0.0).sum(axis=0)
    err = np.abs(got - ref)
    assert err.max() <= 1e-2 * np.abs(ref).max(), err.max()
    assert np.linalg.norm(got - ref) <= 4e-3 * np.linalg.norm(ref)


def _fold_arg(S, pos_t, origin, inv_r, Wt=None):
    f = S.NnFold()
    f.pos = pos_t.data_ptr()
    f.Wt = Wt.data_ptr() if Wt is not None else 0
    for k in range(3):
        f.origin[k] = float(origin[k])
    f.inv_r = float(inv_r)
    return f


@pytest.mark.parametrize("reach", [0.0, 12.0])
def test_folded_half_edge_sums_match_fp64(S, reach):
    """Position fold: with B_j + Wc3 P_j (P = (x - origin) / r, rounded to half
    on the host), the rows' distances and the fold passed, the edge kernel
    subtracts Wc3 P_i from A_i and multiplies only the distance per edge.  The
    sums equal the FP64 sums of the unfolded formula to the unfolded kernel's
    tolerance, also with the origin moved so that |P| reaches 16 radii (the
    limit the surrogate's host code keeps); rows cut to i % 71 entries (empty, tails of every length)."""
    sur, sc, pos, vh, cls, tgt, mask = _small_problem(S, seed=3)
    n = sc.n
    ci, deg, cap, dist = sur.graph_rows(pos, row_cap=16, dist=True)
    D = np.minimum(deg[:n].cpu().numpy(), np.arange(n) % 71).astype(np.int32)
    degc = torch.tensor(D, device="cuda:0")
    CI = ci.cpu().numpy().reshape(-1)[:n * cap].reshape(n, cap)
    rng = np.random.default_rng(3)
    ab = (rng.standard_normal((n, 256)) * 0.7).astype(np.float16)
    Wc = (rng.standard_normal((128, 4)) * 0.8).astype(np.float32)
    bc = (rng.standard_normal(128) * 0.3).astype(np.float32)
    P = pos.cpu().numpy().astype(np.float64)
    r = sur.radius
    origin = 0.5 * (P.min(0) + P.max(0)) - np.array([reach * r, 0.0, 0.0])
    Pn = (P - origin) / r
    assert np.abs(Pn).max() > (15.0 if reach else 4.0)
    abf = ab.copy()
    abf[:, 128:] = (ab[:, 128:].astype(np.float64) + Pn @ Wc[:, :3].T.astype(np.float64)).astype(np.float16)
    lib = S.nnlib()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    abd = torch.tensor(abf.view(np.int16), device="cuda:0")
    Wcd, bcd = torch.tensor(Wc, device="cuda:0"), torch.tensor(bc, device="cuda:0")
    out = torch.zeros((n, 128), dtype=torch.int16, device="cuda:0")
    fd = _fold_arg(S, pos, origin, 1.0 / r)
    assert lib.dmph_nn_edge_fwd_f16(_p(abd), None, _p(ci), _p(dist), _p(Wcd), _p(bcd), _p(out), 128, n,
                                    _p(degc), cap, None, C.byref(fd), s) == 0
    # the fold needs rows (distance pairs): CSR is refused
    assert lib.dmph_nn_edge_fwd_f16(_p(abd), _p(degc), _p(ci), _p(dist), _p(Wcd), _p(bcd), _p(out), 128, n,
                                    None, 0, None, C.byref(fd), s) == -1
    got = _bf16_float(out.cpu().numpy().view(np.uint16))
    A, B = ab[:, :128].astype(np.float64), ab[:, 128:].astype(np.float64)
    ref = np.zeros((n, 128))
    assert D.min() == 0 and D.max() == 70
    for i in range(n):
        J = CI[i, :D[i]]
        d = P[J] - P[i]
        e = np.concatenate([d, np.linalg.norm(d, axis=1, keepdims=True)], axis=1) / r
        ref[i] = np.maximum(A[i] + bc + B[J] + e @ Wc.T.astype(np.float64), 0.0).sum(axis=0)
    err = np.abs(got - ref)
    assert err.max() <= 1e-2 * np.abs(ref).max(), err.max()
    assert np.linalg.norm(got - ref) <= 4e-3 * np.linalg.norm(ref)


def _check_fold(ab_u, ab_f, WP):
    """A half bit for bit; B half = unfolded + Wc3 P to the half rounding of both"""
    assert torch.equal(ab_u[:, :128], ab_f[:, :128])
    u = ab_u[:, 128:].cpu().numpy().view(np.float16).astype(np.float64)
    f = ab_f[:, 128:].cpu().numpy().view(np.float16).astype(np.float64)
    assert np.isfinite(f).all()
    assert (np.abs(f - (u + WP)) <= 2.0 ** -10 * (np.abs(u) + np.abs(WP)) + 1e-4).all()
    assert np.abs(WP).max() > 1.0   # the term is there


def test_node_and_encode_fold_add_the_position_term(S):
    """dmph_nn_node and dmph_nn_encode with a fold write the same h' and A
    half as without, and B + Wc3 (x - origin) / r of each row (Wt = Wc3
    transposed), to half rounding."""
    lib = S.nnlib()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(12)
    Wc = (rng.standard_normal((128, 4)) * 0.8).astype(np.float32)
    Wt = torch.tensor(np.ascontiguousarray(Wc[:, :3].T), device="cuda:0")
    # node update
    M = 3001
    HSd = _dev_bf16(rng.standard_normal((M, 256)).astype(np.float32))
    W1d = _dev_bf16((rng.standard_normal((128, 256)) / 16).astype(np.float32))
    W2d = _dev_bf16((rng.standard_normal((128, 128)) / 11).astype(np.float32))
    Wabd = _dev_bf16((rng.standard_normal((256, 128)) / 11).astype(np.float32))
    b1, b2 = (torch.tensor(rng.standard_normal(128).astype(np.float32), device="cuda:0") for _ in range(2))
    P = rng.uniform(0.0, 1.0, (M, 3)).astype(np.float32)
    pos = torch.tensor(P, device="cuda:0")
    origin, r = np.array([0.5, 0.4, 0.6]), 0.06
    fd = _fold_arg(S, pos, origin, 1.0 / r, Wt)
    outs = []
    for f in (None, C.byref(fd)):
        h = torch.zeros((M, 256), dtype=torch.int16, device="cuda:0")
        ab = torch.full((M, 256), -1, dtype=torch.int16, device="cuda:0")
        assert lib.dmph_nn_node(_p(HSd), 256, _p(W1d), _p(b1), _p(W2d), _p(b2), _p(Wabd), _p(h), 256, _p(ab), M, f,
                                s) == 0
        outs.append((h, ab))
    assert torch.equal(outs[0][0], outs[1][0])
    _check_fold(outs[0][1], outs[1][1], ((P.astype(np.float64) - origin) / r) @ Wc[:, :3].T.astype(np.float64))
    # encoder
    sur, sc, pos, vh, cls, tgt, mask = _small_problem(S, seed=5)
    n = sc.n
    L0 = sur.rounds[0]["ab"]
    Pn = pos.cpu().numpy().astype(np.float64)
    origin = 0.5 * (Pn.min(0) + Pn.max(0))
    fd = _fold_arg(S, pos, origin, 1.0 / sur.radius, Wt)
    outs = []
    for f in (None, C.byref(fd)):
        h0 = torch.zeros((n, 256), dtype=torch.int16, device="cuda:0")
        ab = torch.zeros((n, 256), dtype=torch.int16, device="cuda:0")
        assert lib.dmph_nn_encode(_p(vh), n, _p(cls), _p(pos), C.byref(sur._feat()), _p(sur.enc1.wb),
                                  _p(sur.enc1.b), _p(sur.enc2.wb), _p(sur.enc2.b), _p(L0.wb), _p(h0), 256, _p(ab), f,
                                  s) == 0
        outs.append((h0, ab))
    assert torch.equal(outs[0][0], outs[1][0])
    _check_fold(outs[0][1], outs[1][1], ((Pn - origin) / sur.radius) @ Wc[:, :3].T.astype(np.float64))


def test_inference_with_position_fold_matches_fp64_oracle(S):
    """The fused inference path (encoder, node kernels, fixed-capacity rows)
    with the position fold active: outputs within 5% (norm) of the FP64
    oracle and within 1% of the same path without the fold."""
    sur, sc, pos, vh, cls, tgt, mask = _small_problem(S, seed=6)
    n = sc.n
    rp, ci, nnz = sur.graph(pos)
    X0 = _host_bf16(sur.features(vh, cls, pos, train=False)[:n])
    yr, _ = NO.forward(sur.state_dict(), X0, pos.cpu().numpy().astype(np.float64), rp.cpu().numpy(),
                       ci.cpu().numpy()[:nnz], sur.radius)
    ys = []
    for on in (True, False):
        sur.pos_fold = on
        ci2, deg, cap, eps2 = sur.graph_rows(pos)
        assert (sur._fold_args(pos, n) is not None) == on
        ys.append(sur.forward(None, pos, None, ci2, n, train=False, eps=eps2, rows=(deg, cap),
                              feats=(vh, cls), order=sur._graph_ws["order"]).cpu().numpy().copy())
    assert np.isfinite(ys[0]).all()
    assert np.linalg.norm(ys[0] - yr) / np.linalg.norm(yr) < 0.05
    assert np.linalg.norm(ys[0] - ys[1]) / np.linalg.norm(ys[1]) < 0.01


def test_inference_forward_matches_fp64_oracle(S):
    """Inference folds the aggregation's linear layer into the node layer (one
    GEMM fewer per round): outputs within 5% (norm) of the FP64 oracle, also
    after an Adam step has changed the weights the fold was built from."""
    sur, sc, pos, vh, cls, tgt, mask = _small_problem(S, seed=2)
    n = sc.n
    rp, ci, nnz = sur.graph(pos)
    x0 = sur.features(vh, cls, pos)
    count = int(mask.sum().item())
    X0 = _host_bf16(x0[:n])
    prev = None
    for it in range(2):
        y = sur.forward(x0, pos, rp, ci, n, train=False).clone()
        yr, _ = NO.forward(sur.state_dict(), X0, pos.cpu().numpy().astype(np.float64), rp.cpu().numpy(),
                           ci.cpu().numpy()[:nnz], sur.radius)
        assert torch.isfinite(y).all()
        ey = np.linalg.norm(y.cpu().numpy() - yr) / np.linalg.norm(yr)
        assert ey < 0.05, (it, ey)
        if prev is not None:   # the step moved the output well beyond the tolerance: a stale fold would fail
            assert np.linalg.norm(yr - prev) > 0.2 * np.linalg.norm(prev)
        prev = yr
        if it == 0:
            sur.forward(x0, pos, rp, ci, n)
            sur.loss_and_backward(x0, pos, rp, ci, n, tgt, mask, count)
            sur.adam(2e-3)


def test_training_reduces_loss(S):
    """Adam on a target the inputs determine (twice the newest velocity): the
    loss falls at least tenfold in 150 steps."""
    sur, sc, pos, vh, cls, tgt, mask = _small_problem(S, seed=1)
    n = sc.n
    tgt = (2.0 * vh[0]).contiguous()
    rp, ci, nnz = sur.graph(pos)
    x0 = sur.features(vh, cls, pos)
    count = int(mask.sum().item())
    losses = []
    for it in range(150):
        sur.forward(x0, pos, rp, ci, n)
        losses.append(sur.loss_and_backward(x0, pos, rp, ci, n, tgt, mask, count))
        sur.adam(1e-3)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])


def test_half_edge_sums_rows_past_2_22(S):
    """The inference edge sums' buffer-descriptor path (taken while N x 512
    bytes of [A | B] rows fit 4 GB) with rows that reference nodes j >= 2^22,
    whose row offset j x 512 no longer fits a signed int: N = 5,000,000 nodes,
    128 rows (half of them at i >= 2^22) of 40 partners each, drawn from both
    sides of 2^22; the sums against FP64 sums of the same half inputs."""
    n = 5_000_000
    assert n * 512 <= 0xFFFFFFFF - 1024 and n > (1 << 22)
    g = torch.Generator(device="cuda:0")
    g.manual_seed(5)
    abd = (torch.randn((n, 256), generator=g, device="cuda:0", dtype=torch.float32) * 0.7).to(torch.float16)
    pos = torch.rand((n, 3), generator=g, device="cuda:0", dtype=torch.float32) * 0.05
    rng = np.random.default_rng(5)
    rows = np.concatenate([rng.choice(1 << 22, 64, replace=False), (1 << 22) + rng.choice(n - (1 << 22), 64, replace=False)])
    lens = np.zeros(n, np.int64)
    lens[rows] = 40
    RP = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    CI = np.zeros(int(RP[-1]), np.int32)
    for i in rows:
        CI[RP[i]:RP[i + 1]] = np.sort(np.concatenate([rng.integers(0, 1 << 22, 10), rng.integers(1 << 22, n, 30)]))
    rp, ci = torch.tensor(RP, device="cuda:0"), torch.tensor(CI, device="cuda:0")
    Wc = (rng.standard_normal((128, 4)) * 0.8).astype(np.float32)
    bc = (rng.standard_normal(128) * 0.3).astype(np.float32)
    Wcd, bcd = torch.tensor(Wc, device="cuda:0"), torch.tensor(bc, device="cuda:0")
    lib = S.nnlib()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    radius = 0.03
    eps = torch.empty(2 * ci.numel(), dtype=torch.int32, device="cuda:0")
    out = torch.zeros((n, 128), dtype=torch.int16, device="cuda:0")
    assert lib.dmph_nn_edge_eps(_p(pos), _p(rp), _p(ci), _p(eps), n, C.c_float(radius), s) == 0
    assert lib.dmph_nn_edge_fwd_f16(_p(abd), _p(rp), _p(ci), _p(eps), _p(Wcd), _p(bcd), _p(out), 128, n,
                                    None, 0, None, None, s) == 0
    torch.cuda.synchronize()
    need = np.unique(np.concatenate([rows, CI]))   # the nodes in use, compacted
    idx = torch.tensor(need, device="cuda:0")
    AB = abd[idx].cpu().numpy()
    P = pos[idx].cpu().numpy().astype(np.float64)
    at = lambda v: np.searchsorted(need, v)
    got = _bf16_float(out[torch.tensor(rows, device="cuda:0")].cpu().numpy().view(np.uint16))
    ref = np.zeros((len(rows), 128))
    for k, i in enumerate(rows):
        J = at(CI[RP[i]:RP[i + 1]])
        d = P[J] - P[at(i)]
        e = np.concatenate([d, np.linalg.norm(d, axis=1, keepdims=True)], axis=1) / radius
        # the kernel reads the edge features as halves: round them the same way
        e = e.astype(np.float16).astype(np.float64)
        A, B = AB[at(i), :128].astype(np.float64), AB[J, 128:].astype(np.float64)
        ref[k] = np.maximum(A + bc + B + e @ Wc.T.astype(np.float64), 0.0).sum(axis=0)
    assert int(out.count_nonzero()) > 0
    err = np.abs(got - ref)
    assert err.max() <= 1e-2 * np.abs(ref).max(), err.max()
    assert np.linalg.norm(got - ref) <= 4e-3 * np.linalg.norm(ref)
