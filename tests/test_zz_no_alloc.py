"""The allocation-free sift pass (icm_sift_pass_static) against the
allocating one (icm_sift_pass), through the C ABI: two calls in a row, the
second with a smaller input, bit for bit in both precisions, and the size
limit refused on the host."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def _call(icm, fn, h, hn, nk):
    import torch
    st = C.c_void_p(torch.cuda.current_stream(h.device).cuda_stream)
    return fn(icm.DTYPE_F64 if h.dtype == torch.float64 else icm.DTYPE_F32,
              C.c_void_p(h.data_ptr()), h.shape[1], h.shape[0], C.c_void_p(hn.data_ptr()),
              C.c_void_p(nk.data_ptr()), st)


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_static_sift_pass_matches_allocating_one(icm, dev, dt):
    import importlib
    import torch
    sig = importlib.import_module("ceemdan-gpu_amd.signals")
    dtype = torch.float64 if dt == "f64" else torch.float32
    L = icm.lib()
    for I, N in ((24, 70000), (5, 3001)):   # the second call is smaller
        print(f"case: static sift pass {dt} I={I} N={N}", flush=True)
        h = torch.stack([torch.tensor(sig.eeg_like(N, seed=700 + i), dtype=dtype)
                         for i in range(I)]).to(dev)
        a, b = torch.empty_like(h), torch.full_like(h, float("nan"))
        ka = torch.empty((I, 2), dtype=torch.int32, device=dev)
        kb = torch.full((I, 2), -1, dtype=torch.int32, device=dev)
        assert _call(icm, L.icm_sift_pass, h, a, ka) == 0
        assert _call(icm, L.icm_sift_pass_static, h, b, kb) == 0
        torch.cuda.synchronize()
        assert torch.equal(a, b), (I, N)
        assert torch.equal(ka, kb), (I, N)


def test_static_sift_pass_limits(icm, dev):
    import torch
    h = torch.zeros((4097, 8), dtype=torch.float64, device=dev)
    hn = torch.empty_like(h)
    nk = torch.empty((4097, 2), dtype=torch.int32, device=dev)
    assert _call(icm, icm.lib().icm_sift_pass_static, h, hn, nk) == -1
    h = torch.zeros((2, (1 << 20) + 1), dtype=torch.float32, device=dev)
    hn = torch.empty_like(h)
    nk = torch.empty((2, 2), dtype=torch.int32, device=dev)
    assert _call(icm, icm.lib().icm_sift_pass_static, h, hn, nk) == -1
