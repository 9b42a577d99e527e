"""GPU parity of the HIP path against the FP64 CPU oracle.

FP64: bit-identical results are expected (same arithmetic order, no FMA
contraction on either side); the assertions use exact equality for knot
counts and a 1e-12 relative bound for samples (the required bar for FP64
IMFs is 1e-9).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _eeg(name, n=None):
    x = np.fromfile(os.path.join(GOLD, name), dtype="<f4").astype(np.float64)
    return x if n is None else x[:n]


def _sig():
    import importlib
    return importlib.import_module("ceemdan-gpu_amd.signals")


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_noise_matches_oracle(icm, orc, dev):
    import torch
    print("case: noise I=7 N=1001 seed=1", flush=True)
    w = icm.noise(7, 1001, seed=1, dtype=torch.float64, device=dev).cpu().numpy()
    ref = orc.noise(1, 7, 1001)
    # identical Philox bits; Box-Muller's log/sin/cos come from different libms
    assert np.max(np.abs(w - ref)) < 1e-13
    w32 = icm.noise(7, 1001, seed=1, dtype=torch.float32, device=dev).cpu().numpy()
    assert np.array_equal(w32, ref.astype(np.float32)) or np.max(np.abs(w32 - ref)) < 1e-6


@pytest.mark.parametrize("case", ["white", "eeg", "sine", "short", "plateau"])
def test_sift_pass_fp64_bitexact(icm, orc, dev, case):
    import torch
    rng = np.random.default_rng(11)
    if case == "white":
        h = rng.standard_normal((5, 3000))
    elif case == "eeg":
        h = np.stack([_eeg("eegSampleDataCH4.f32", 30504)[i * 5000:(i + 1) * 5000 + 77]
                      for i in range(4)])
    elif case == "sine":
        n = np.arange(2600)
        h = np.stack([np.sin(2 * np.pi * n / p) + 0.01 * rng.standard_normal(n.size)
                      for p in (7.3, 40.0, 900.0)])
    elif case == "short":
        h = rng.standard_normal((6, 9))
    else:
        h = np.round(rng.standard_normal((3, 2000)) * 2) / 2   # many equal neighbours
    print(f"case: sift_pass fp64 {case} shape={h.shape}", flush=True)
    hd = torch.tensor(h, dtype=torch.float64, device=dev)
    hn, nk = icm.sift_pass(hd)
    hn = hn.cpu().numpy()
    nk = nk.cpu().numpy()
    for i in range(h.shape[0]):
        ref, nmax, nmin, _, _ = orc.sift_pass(h[i])
        assert (nk[i, 0], nk[i, 1]) == (nmax, nmin), i
        assert np.array_equal(hn[i], ref), (i, relerr(hn[i], ref))


def test_sift_pass_long_series(icm, orc, dev):
    """Many tiles and several solve windows per series (EEG CH1, 244032 samples)."""
    import torch
    x = _eeg("eegSampleDataCH1_244032.f32")
    h = np.stack([x, x[::-1].copy()])
    print(f"case: sift_pass fp64 long shape={h.shape}", flush=True)
    hn, nk = icm.sift_pass(torch.tensor(h, dtype=torch.float64, device=dev))
    hn = hn.cpu().numpy()
    nk = nk.cpu().numpy()
    for i in range(2):
        ref, nmax, nmin, _, _ = orc.sift_pass(h[i])
        assert (nk[i, 0], nk[i, 1]) == (nmax, nmin)
        assert np.array_equal(hn[i], ref), relerr(hn[i], ref)


def test_iceemdan_fp64_small_bitexact(icm, orc, dev):
    """Three-tone signal, N=4096, I=12, M=6, FIXED S=10, noise from the GPU."""
    import torch
    N, I, M, S = 4096, 12, 6, 10
    x = _sig().three_tone(N, seed=0)
    print(f"case: iceemdan fp64 N={N} I={I} M={M} S={S}", flush=True)
    w = icm.noise(I, N, seed=1, dtype=torch.float64, device=dev)
    plan = icm.Plan([N], I, M, dtype=torch.float64, max_sifts=S, use_graph=False, device=dev)
    plan.enable_trace(S)
    out = plan.forward([torch.tensor(x, dtype=torch.float64, device=dev)], noise=w)[0]
    got = out.cpu().numpy()
    ref = orc.iceemdan(x, w.cpu().numpy(), M, S_or_cap=S, trace_passes=S)
    tr_sig = plan.trace(icm.PHASE_SIGNAL)     # [M-1][S][I][2]
    tr_noise = plan.trace(icm.PHASE_NOISE)
    assert np.array_equal(tr_noise, ref["trace"][0])
    assert np.array_equal(tr_sig, ref["trace"][1])
    assert np.array_equal(got, ref["imfs"]), relerr(got, ref["imfs"])
    plan.close()


def test_golden_fixtures_on_gpu(icm, dev):
    """The HIP path reproduces the committed oracle outputs
    (tests/golden/oracle_fixtures.npz) from the fixtures' own noise: FP64
    modes and Rilling stop counts bit for bit, the gradient to 1e-9."""
    import torch
    data = np.load(os.path.join(GOLD, "oracle_fixtures.npz"), allow_pickle=False)
    x = _eeg("eegSampleDataCH4.f32")
    print("case: golden noise", flush=True)
    w = icm.noise(3, 37, seed=1234, dtype=torch.float64, device=dev).cpu().numpy()
    assert np.max(np.abs(w - data["noise/w"])) < 1e-13

    print("case: golden sift_eeg", flush=True)
    hn, nk = icm.sift_pass(torch.tensor(x[:1500], device=dev))
    assert np.array_equal(hn[0].cpu().numpy(), data["sift_eeg/h_new"])
    assert tuple(nk[0].cpu().numpy()) == tuple(data["sift_eeg/counts"][:2])

    print("case: golden fixed_eeg", flush=True)
    w = torch.tensor(data["fixed_eeg/noise"], device=dev)
    got = icm.iceemdan(torch.tensor(x[:2000], device=dev), 6, 5, max_sifts=10, noise=w)
    assert np.array_equal(got.cpu().numpy(), data["fixed_eeg/imfs"])

    print("case: golden rilling_two_tone", flush=True)
    xt = data["rilling_two_tone/x"]     # stored: libm sin differs across hosts
    w = torch.tensor(data["rilling_two_tone/noise"], device=dev)
    plan = icm.Plan([1000], 8, 3, dtype=torch.float64, stop_mode=icm.STOP_RILLING, max_sifts=100,
                    noise_max_sifts=2000, use_graph=False, device=dev)
    plan.enable_trace(1)
    got = plan.forward([torch.tensor(xt, device=dev)], noise=w)[0].cpu().numpy()
    stops = np.stack([plan.stops(icm.PHASE_NOISE), plan.stops(icm.PHASE_SIGNAL)])
    plan.close()
    assert np.array_equal(stops, data["rilling_two_tone/stops"])
    assert np.array_equal(got, data["rilling_two_tone/imfs"])

    print("case: golden grad_eeg", flush=True)
    w = torch.tensor(data["grad_eeg/noise"], device=dev)
    plan = icm.Plan([800], 4, 4, dtype=torch.float64, max_sifts=6, use_graph=False, device=dev)
    plan.forward([torch.tensor(x[:800], device=dev)], noise=w)
    G = np.cos(np.arange(4 * 800).reshape(4, 800) / 17.0)
    g = plan.backward([torch.tensor(G, device=dev)])[0].cpu().numpy()
    plan.close()
    assert relerr(g, data["grad_eeg/grad"]) < 1e-9


@pytest.mark.parametrize("stop_mode", [0, 1])
def test_not_a_knot_fp64_bitexact(icm, orc, dev, stop_mode):
    """Not-a-knot spline ends: the HIP path equals the oracle bit for bit
    (FIXED and RILLING); the adjoint is natural-ends only."""
    import torch
    N, I, M = 3000, 6, 5
    x = _eeg("eegSampleDataCH4.f32", 30504)[7000:7000 + N]
    print(f"case: not-a-knot stop_mode={stop_mode}", flush=True)
    w = icm.noise(I, N, seed=11, dtype=torch.float64, device=dev)
    S = 10 if stop_mode == 0 else 60
    plan = icm.Plan([N], I, M, dtype=torch.float64, stop_mode=stop_mode, max_sifts=S,
                    noise_max_sifts=S if stop_mode == 0 else 300, use_graph=False, device=dev,
                    boundary=icm.BOUNDARY_NOT_A_KNOT)
    got = plan.forward([torch.tensor(x, device=dev)], noise=w)[0].cpu().numpy()
    ref = orc.iceemdan(x, w.cpu().numpy(), M, stop_mode=stop_mode, S_or_cap=S,
                       noise_cap=S if stop_mode == 0 else 300, bc=orc.BC_NOT_A_KNOT)["imfs"]
    assert np.array_equal(got, ref), relerr(got, ref)
    if stop_mode == 0:
        with pytest.raises(icm.ICMError):
            plan.backward([torch.zeros((M, N), dtype=torch.float64, device=dev)])
    plan.close()


def test_long_series_bitexact(icm, orc, dev):
    """A series of more than 1024 tiles (300,000 samples) takes the chunked
    knot-prefix kernel; FP64 result equal to the oracle."""
    import torch
    N, I, M, S = 300000, 2, 3, 3
    x = _sig().eeg_like(N, seed=77)
    print(f"case: long series N={N}", flush=True)
    w = icm.noise(I, N, seed=5, dtype=torch.float64, device=dev)
    got = icm.iceemdan(torch.tensor(x, device=dev), I, M, max_sifts=S, noise=w).cpu().numpy()
    ref = orc.iceemdan(x, w.cpu().numpy(), M, S_or_cap=S)["imfs"]
    assert np.array_equal(got, ref), relerr(got, ref)


@pytest.mark.parametrize("sparse_stage,dense_max,stop_mode",
                         [(0, 1 << 30, 0), (0, 1 << 30, 1), (0, 0, 0), (1, -1, 0)])
def test_small_knot_tables_bitexact(icm, orc, dev, monkeypatch, sparse_stage, dense_max,
                                    stop_mode):
    """Sift passes with small knot tables: tiles with more knots than they
    hold go to a second launch with full tables, and the next tile of their
    segment evaluates h_new left of itself from its own knots; a pass whose
    input has more such tiles than tolerated runs with full tables in the
    second launch.  Small tables from stage 0 on (white-noise first stages:
    most tiles overflow) with every overflow tolerated, with none tolerated,
    and the default; segments of 8 tiles.  FP64 modes equal the oracle bit
    for bit."""
    import torch
    monkeypatch.setenv("ICM_SPARSE_STAGE", str(sparse_stage))
    monkeypatch.setenv("ICM_SPARSE_DENSE_MAX", str(dense_max))
    monkeypatch.setenv("ICM_SEGMENT_TILES", "8")
    N, I, M = 20000, 6, 5
    x = _sig().eeg_like(N, seed=3)
    print(f"case: small knot tables sparse_stage={sparse_stage} dense_max={dense_max} "
          f"stop_mode={stop_mode}", flush=True)
    w = icm.noise(I, N, seed=21, dtype=torch.float64, device=dev)
    S = 10 if stop_mode == 0 else 40
    plan = icm.Plan([N], I, M, dtype=torch.float64, stop_mode=stop_mode, max_sifts=S,
                    noise_max_sifts=S if stop_mode == 0 else 200, use_graph=False, device=dev)
    got = plan.forward([torch.tensor(x, device=dev)], noise=w)[0].cpu().numpy()
    plan.close()
    ref = orc.iceemdan(x, w.cpu().numpy(), M, stop_mode=stop_mode, S_or_cap=S,
                       noise_cap=S if stop_mode == 0 else 200)["imfs"]
    assert np.array_equal(got, ref), relerr(got, ref)
