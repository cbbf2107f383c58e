"""dmph_nn_mse_static (the loss with its partials in the library's fixed
device array, no scratch argument, no allocation) against dmph_nn_mse (caller
scratch) on the same inputs: loss and gradient bit for bit, on two calls in a
row with the second smaller, and the size limit reported as -4."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MAX_N = 65536 * 256   # DMPH_NN_MSE_STATIC_MAX_N in include/dem_mph_nn.h


def _p(t):
    return C.c_void_p(t.data_ptr())


def _case(N, count, seed):
    rng = np.random.default_rng(seed)
    Y = torch.tensor(rng.standard_normal((N, 3)).astype(np.float32), device="cuda:0")
    T = torch.tensor((3.0 * rng.standard_normal((N, 3))).astype(np.float32), device="cuda:0")
    mask = torch.tensor(rng.integers(0, 6, N).astype(np.int32), device="cuda:0")
    return Y, T, mask


def _both(lib, N, count, clip, seed):
    Y, T, mask = _case(N, count, seed)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    dY_a = torch.full((N, 3), float("nan"), device="cuda:0")
    dY_b = torch.full((N, 3), float("nan"), device="cuda:0")
    loss_a = torch.full((1,), float("nan"), device="cuda:0")
    loss_b = torch.full((1,), float("nan"), device="cuda:0")
    ws = torch.empty(int(lib.dmph_nn_mse_scratch(N)) // 4 + 1, dtype=torch.float32, device="cuda:0")
    rc = lib.dmph_nn_mse(_p(Y), _p(T), _p(mask), N, count, C.c_float(clip), _p(dY_a), _p(loss_a), _p(ws), s)
    assert rc == 0
    rc = lib.dmph_nn_mse_static(_p(Y), _p(T), _p(mask), N, count, C.c_float(clip), _p(dY_b), _p(loss_b), s)
    assert rc == 0
    torch.cuda.synchronize()
    return (loss_a.cpu().numpy(), dY_a.cpu().numpy()), (loss_b.cpu().numpy(), dY_b.cpu().numpy()), (Y, T, mask)


def test_mse_static_matches_scratch_variant():
    from dem_mph_amd import surrogate as SG
    lib = SG.nnlib()
    # two calls in a row, the second smaller: the second must not see the
    # first call's partials left in the shared array
    for N, count, seed in ((1_000_003, 700_000, 1), (5_000, 4_321, 2)):
        print("mse_static N=%d" % N, flush=True)
        (la, da), (lb, db), (Y, T, mask) = _both(lib, N, count, 2.5, seed)
        assert np.isfinite(la).all() and np.isfinite(da).all()
        np.testing.assert_array_equal(lb, la)
        np.testing.assert_array_equal(db, da)
        # and both equal the definition (FP64 on the host) to FP32 rounding
        y, t, m = Y.cpu().numpy().astype(np.float64), T.cpu().numpy().astype(np.float64), mask.cpu().numpy()
        d = np.where(m[:, None] > 0, y - np.clip(t, -2.5, 2.5), 0.0)
        ref = float((m[:, None] * d * d).sum() / count)
        assert abs(float(lb[0]) - ref) <= 1e-5 * ref


def test_mse_static_rejects_sizes_beyond_its_array():
    from dem_mph_amd import surrogate as SG
    lib = SG.nnlib()
    null = C.c_void_p(0)
    # rejected before any launch, so no device pointer is touched
    rc = lib.dmph_nn_mse_static(null, null, null, MAX_N + 1, 1, C.c_float(1.0), null, null, null)
    assert rc == -4
    rc = lib.dmph_nn_mse_static(null, null, null, 0, 1, C.c_float(1.0), null, null, null)
    assert rc == -1
