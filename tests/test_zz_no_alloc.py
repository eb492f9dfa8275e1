"""The predictor's loss/gradient entry point that allocates nothing
(NESTGPU_AMD_PredictorLossGradNoAlloc, device buffers in and out, partial
sums in the library's fixed array) against the allocating entry point
(NESTGPU_AMD_PredictorLossGrad, host buffers) on the same batch: equal bit
for bit, on two calls in a row, the second with a smaller batch (the array
keeps the first call's larger rows, which the second must not read)."""
import ctypes

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

N_TRAIN = 1377
MAX_N = 16384


def _params(rng):
    P = np.zeros(1393, np.float32)
    P[0:8] = rng.normal(0, 1, 8)
    P[8:16] = rng.uniform(0.5, 2.0, 8)
    P[16:] = rng.normal(0, 0.3, 1393 - 16)
    P[-1] = -1.0
    return P


def _batch(rng, n):
    raw = rng.normal(0, 2.0, (n, 8)).astype(np.float32)
    target = rng.normal(-2.0, 1.0, n).astype(np.float32)
    cens = (rng.random(n) < 0.4).astype(np.uint8)
    return raw, target, cens


def test_loss_grad_noalloc_matches_allocating_entry():
    if not gpu_available():
        pytest.skip("no GPU")
    import torch
    import nestgpu
    nestgpu.SetVerbosityLevel(0)
    lib = nestgpu._nestgpu
    f = lib.NESTGPU_AMD_PredictorLossGradNoAlloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] + [ctypes.c_void_p] * 3
    rng = np.random.default_rng(5)
    nestgpu.SetPredictorParams(_params(rng))
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    try:
        for n in (MAX_N, 1000):     # full array, then a smaller batch
            print("no-alloc loss/grad: n = %d" % n, flush=True)
            raw, target, cens = _batch(rng, n)
            loss_ref, grad_ref = nestgpu.PredictorLossGrad(raw, target, cens)
            d_raw = torch.from_numpy(raw).to(dev)
            d_t = torch.from_numpy(target).to(dev)
            d_c = torch.from_numpy(cens).to(dev)
            d_loss = torch.full((1,), float("nan"), dtype=torch.float32, device=dev)
            d_grad = torch.full((N_TRAIN,), float("nan"), dtype=torch.float32, device=dev)
            torch.cuda.synchronize()
            rc = f(d_raw.data_ptr(), d_t.data_ptr(), d_c.data_ptr(), n, d_loss.data_ptr(),
                   d_grad.data_ptr(), ctypes.c_void_p(stream.cuda_stream))
            assert rc == 0 and nestgpu.GetErrorCode() == 0, nestgpu.GetErrorMessage()
            torch.cuda.synchronize()
            loss = float(d_loss.cpu()[0])
            grad = d_grad.cpu().numpy()
            assert np.isfinite(loss_ref) and loss_ref > 0.0
            assert loss == loss_ref, (n, loss, loss_ref)
            np.testing.assert_array_equal(grad, grad_ref)
        # a batch beyond the array is refused before anything is launched
        # (the library's error convention: error code 2 and a message)
        raw, target, cens = _batch(rng, 64)
        d_raw = torch.from_numpy(raw).to(dev)
        d_t = torch.from_numpy(target).to(dev)
        d_c = torch.from_numpy(cens).to(dev)
        d_loss = torch.zeros(1, dtype=torch.float32, device=dev)
        d_grad = torch.zeros(N_TRAIN, dtype=torch.float32, device=dev)
        rc = f(d_raw.data_ptr(), d_t.data_ptr(), d_c.data_ptr(), MAX_N + 1, d_loss.data_ptr(),
               d_grad.data_ptr(), ctypes.c_void_p(stream.cuda_stream))
        assert nestgpu.GetErrorCode() == 2
        assert b"16384" in nestgpu.GetErrorMessage()
        assert float(d_loss.cpu()[0]) == 0.0     # outputs untouched
    finally:
        nestgpu.SetStepPredictorMode(0)
