"""GPU tests of the plan features: FP32 accuracy against the FP64 oracle,
Rilling stop rule, drop-in entry points, batched channels, hipGraph replay,
statistics kernels, determinism and full-size invariants."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _sig():
    import importlib
    return importlib.import_module("ceemdan-gpu_amd.signals")


def _eeg(name):
    return np.fromfile(os.path.join(GOLD, name), dtype="<f4").astype(np.float64)


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_fp32_sift_pass_close_to_fp64_oracle(icm, orc, dev):
    import torch
    x = _eeg("eegSampleDataCH4.f32")[:20000]
    h = np.stack([x, x[::-1].copy()])
    print("case: sift_pass fp32 eeg 2x20000", flush=True)
    hn, nk = icm.sift_pass(torch.tensor(h, dtype=torch.float32, device=dev))
    hn = hn.cpu().numpy().astype(np.float64)
    nk = nk.cpu().numpy()
    for i in range(2):
        # the oracle is fed the same float32-rounded samples
        ref, nmax, nmin, _, _ = orc.sift_pass(h[i].astype(np.float32).astype(np.float64))
        assert (nk[i, 0], nk[i, 1]) == (nmax, nmin)
        e = relerr(hn[i], ref)
        print(f"fp32 sift pass {i}: max rel err {e:.2e}", flush=True)
        assert e < 1e-4


def test_fp32_iceemdan_vs_fp64_oracle(icm, orc, dev):
    """Per-mode max relative error of the FP32 path against the FP64 oracle
    fed the same float32 input and noise: measured 1.1e-6, 2.6e-6, 1.1e-6,
    8.2e-7, 1.8e-5 and 2.0e-4 (residue).  The first four rows carry FP32
    rounding only and are held to 1e-5; the last two carry the few-knot late
    stages, where a single extremum found differently by the two precisions
    moves an envelope locally (the residue is also small, which inflates its
    relative error), and are held to 1e-3."""
    import torch
    N, I, M, S = 4096, 16, 6, 10
    x = _sig().three_tone(N, seed=0)
    print(f"case: iceemdan fp32 N={N} I={I} M={M}", flush=True)
    w = icm.noise(I, N, seed=1, dtype=torch.float64, device=dev)
    w32 = w.to(torch.float32)
    out = icm.iceemdan(torch.tensor(x, dtype=torch.float32, device=dev), I, M, max_sifts=S,
                       noise=w32).cpu().numpy().astype(np.float64)
    ref = orc.iceemdan(x.astype(np.float32).astype(np.float64),
                       w32.cpu().numpy().astype(np.float64), M, S_or_cap=S)["imfs"]
    errs = [relerr(out[k], ref[k]) for k in range(M)]
    print("fp32 per-mode max rel err:", " ".join(f"{e:.2e}" for e in errs))
    assert np.all(np.isfinite(out))
    assert max(errs[:4]) < 1e-5
    assert max(errs) < 1e-3


def test_rilling_fp64_matches_oracle(icm, orc, dev):
    """Reference stop rule (cudaICEEMDAN.cu:494-544) with per-series masks:
    same subtraction counts and bit-identical modes."""
    import torch
    N, I, M, cap = 1000, 20, 3, 100
    x = _sig().two_tone(N)
    print(f"case: rilling fp64 two-tone N={N} I={I} M={M}", flush=True)
    w = icm.noise(I, N, seed=1234, dtype=torch.float64, device=dev)
    plan = icm.Plan([N], I, M, dtype=torch.float64, stop_mode=icm.STOP_RILLING, max_sifts=cap,
                    noise_max_sifts=2000, use_graph=False, device=dev)
    plan.enable_trace(4)
    got = plan.forward([torch.tensor(x, dtype=torch.float64, device=dev)], noise=w)[0]
    got = got.cpu().numpy()
    ref = orc.iceemdan(x, w.cpu().numpy(), M, stop_mode=orc.STOP_RILLING, S_or_cap=cap,
                       noise_cap=2000)
    assert np.array_equal(plan.stops(icm.PHASE_NOISE), ref["stops"][0])
    assert np.array_equal(plan.stops(icm.PHASE_SIGNAL), ref["stops"][1])
    assert np.array_equal(got, ref["imfs"]), relerr(got, ref["imfs"])
    plan.close()


def test_dropin_iceemdanD_matches_oracle(icm, orc, dev):
    """iceemdanD with the reference's arguments (sample_synthetic_signal.cu:39-55):
    Rilling rule, noise-phase cap 2000, seed 1234."""
    import torch
    N, I, M, max_iter = 1000, 100, 2, 100
    x = _sig().two_tone(N)
    print("case: dropin iceemdanD two-tone", flush=True)
    d_x = torch.arange(N, dtype=torch.int32, device=dev)
    d_y = torch.tensor(x, dtype=torch.float64, device=dev)
    d_imfs = torch.empty(M * N, dtype=torch.float64, device=dev)
    t = icm.iceemdanD(I, N, M, max_iter, d_x, d_y, d_imfs, 0.2)
    assert t >= 0.0
    w = icm.noise(I, N, seed=1234, dtype=torch.float64, device=dev).cpu().numpy()
    ref = orc.iceemdan(x, w, M, stop_mode=orc.STOP_RILLING, S_or_cap=max_iter, noise_cap=2000)
    got = d_imfs.cpu().numpy().reshape(M, N)
    assert np.array_equal(got, ref["imfs"]), relerr(got, ref["imfs"])


def test_dropin_iceemdanS_and_index_check(icm, orc, dev):
    import torch
    N, I, M = 2000, 20, 4
    x = _eeg("eegSampleDataCH4.f32")[:N]
    print("case: dropin iceemdanS eeg", flush=True)
    d_x = torch.arange(N, dtype=torch.int32, device=dev)
    d_y = torch.tensor(x, dtype=torch.float32, device=dev)
    d_imfs = torch.empty(M * N, dtype=torch.float32, device=dev)
    t = icm.iceemdanS(I, N, M, 50, d_x, d_y, d_imfs, 0.2)
    assert t >= 0.0
    got = d_imfs.cpu().numpy().reshape(M, N).astype(np.float64)
    assert np.all(np.isfinite(got))
    xt = x.astype(np.float32).astype(np.float64)
    xt = xt / np.sqrt(np.var(xt, ddof=1))
    assert np.max(np.abs(got.sum(0) - xt)) < 1e-4
    # d_x must hold the sample numbers (cudaICEEMDAN.cu:423): rejected otherwise
    bad = d_x.flip(0).contiguous()
    assert icm.iceemdanS(I, N, M, 50, bad, d_y, d_imfs, 0.2) < 0


def test_batched_equals_per_signal(icm, dev):
    """Channels of uneven length in one plan give the same bits as one plan
    per channel (noise modes are shared per distinct length)."""
    import torch
    sig = _sig()
    lens = [3000, 4100, 3000, 777]
    xs = [sig.eeg_like(n, seed=10 + c) for c, n in enumerate(lens)]
    I, M, S = 8, 5, 10
    print(f"case: batched fp64 lens={lens}", flush=True)
    xd = [torch.tensor(x, dtype=torch.float64, device=dev) for x in xs]
    plan = icm.Plan(lens, I, M, dtype=torch.float64, max_sifts=S, use_graph=False, device=dev)
    outs = plan.forward(xd)
    plan.close()
    for c, x in enumerate(xd):
        single = icm.iceemdan(x, I, M, max_sifts=S)
        assert torch.equal(outs[c], single), c


def test_graph_replay_matches_eager(icm, dev):
    import torch
    N, I, M, S = 5000, 10, 5, 10
    x = torch.tensor(_sig().eeg_like(N, seed=5), dtype=torch.float64, device=dev)
    print("case: graph replay", flush=True)
    eager = icm.Plan([N], I, M, dtype=torch.float64, max_sifts=S, use_graph=False, device=dev)
    ref = eager.forward([x])[0].clone()
    eager.close()
    g = icm.Plan([N], I, M, dtype=torch.float64, max_sifts=S, use_graph=True, device=dev)
    g.set_autotune(0)   # every forward is a graph replay (tuning forwards run eagerly)
    for k in range(3):
        out = g.forward([x])[0]
        assert torch.equal(out, ref), k
        # unrelated eager work between replays
        junk = torch.randn(1000, device=dev).sum()
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
    g.close()


def test_statistics_kernels(icm, dev):
    import torch
    rng = np.random.default_rng(2)
    for dt in (torch.float64, torch.float32):
        x = rng.standard_normal((3, 10007)) * 2 + 1
        xd = torch.tensor(x, dtype=dt, device=dev)
        means = torch.zeros(3, dtype=dt, device=dev)
        icm.meanS(xd, 10007, 3, means)
        icm.multiplyS(means, means, 1.0 / 10007, 3)
        var = torch.zeros(3, dtype=dt, device=dev)
        icm.varianceS(xd, 10007, 3, means, var)
        icm.multiplyS(var, var, 1.0 / 10006, 3)
        tol = 1e-12 if dt == torch.float64 else 1e-5
        assert np.allclose(means.cpu().numpy(), x.mean(1), rtol=tol, atol=tol)
        assert np.allclose(var.cpu().numpy(), x.var(1, ddof=1), rtol=tol * 10)


def test_deterministic(icm, dev):
    import torch
    N, I, M = 6000, 12, 6
    x = torch.tensor(_sig().eeg_like(N, seed=9), dtype=torch.float64, device=dev)
    a = icm.iceemdan(x, I, M)
    b = icm.iceemdan(x, I, M)
    assert torch.equal(a, b)


def test_full_size_invariants_fp64(icm, dev):
    """One EEG-like channel at the benchmark size (N=244032, I=500, M=16,
    10 sifts): completeness sum(rows) == x/std(x), finite, and a fixed-count
    decomposition from an exactly reproducible run."""
    import torch
    N, I, M = 244032, 500, 16
    x = _sig().eeg_like(N, seed=2)
    print("case: full size fp64 N=244032 I=500 M=16", flush=True)
    xd = torch.tensor(x, dtype=torch.float64, device=dev)
    plan = icm.Plan([N], I, M, dtype=torch.float64, max_sifts=10, use_graph=True, device=dev)
    out = plan.forward([xd])[0].cpu().numpy()
    plan.close()
    xt = x / np.sqrt(np.var(x, ddof=1))
    assert np.all(np.isfinite(out))
    assert np.max(np.abs(out.sum(0) - xt)) < 1e-10
    # energy concentrates in the first modes for pink noise (dyadic filter bank)
    power = (out[:-1] ** 2).mean(1)
    assert power[0] > power[-1]


def test_invalid_arguments_raise(icm, dev):
    import torch
    with pytest.raises(icm.ICMError):
        icm.Plan([3], 10, 4, dtype=torch.float64, device=dev)
    with pytest.raises(icm.ICMError):
        icm.Plan([100], 0, 4, dtype=torch.float64, device=dev)
    plan = icm.Plan([100], 4, 3, dtype=torch.float64, device=dev)
    with pytest.raises(ValueError):
        plan.forward([torch.zeros(99, dtype=torch.float64, device=dev)])
    plan.close()



@pytest.mark.parametrize("bad", [float("nan"), float("inf")])
def test_non_finite_input_matches_oracle(icm, orc, dev, bad):
    """A NaN or an infinity in the input: the forward completes (no fault, no
    hang) and its rows equal the oracle's, NaN for NaN (std(x) is not finite,
    so the standardised input and every row are NaN)."""
    import torch
    N, I, M, S = 3000, 4, 4, 6
    x = _sig().eeg_like(N, seed=77)
    x[N // 3] = bad
    print(f"case: non-finite input {bad}", flush=True)
    w = icm.noise(I, N, seed=1234, dtype=torch.float64, device=dev)
    got = icm.iceemdan(torch.tensor(x, device=dev), I, M, max_sifts=S, noise=w).cpu().numpy()
    ref = orc.iceemdan(x, w.cpu().numpy(), M, S_or_cap=S)["imfs"]
    assert np.array_equal(got, ref, equal_nan=True)
