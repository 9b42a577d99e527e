"""GPU tests of the adjoint (icm_backward): against the FP64 oracle's
gradient, against central differences of the GPU forward, batched channels,
torch.autograd, and the state errors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sig():
    import importlib
    return importlib.import_module("ceemdan-gpu_amd.signals")


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _case(N, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(N) / 1000.0
    return np.sin(2 * np.pi * 5 * t) + 0.5 * np.sin(2 * np.pi * 40 * t) + 0.1 * rng.standard_normal(N)


@pytest.mark.parametrize("N,I,M,S", [(1024, 6, 5, 10), (3000, 16, 6, 4), (700, 3, 3, 1)])
def test_backward_fp64_matches_oracle(icm, orc, dev, N, I, M, S):
    """dL/dx for a random dL/dIMF equals the oracle's gradient (global
    Thomas solve of the transposed system there, the windowed solve here;
    different summation orders): relative max error <= 1e-9."""
    import torch
    x = _case(N, seed=N)
    rng = np.random.default_rng(3)
    G = rng.standard_normal((M, N))
    print(f"case: backward fp64 N={N} I={I} M={M} S={S}", flush=True)
    w = icm.noise(I, N, seed=7, dtype=torch.float64, device=dev)
    plan = icm.Plan([N], I, M, dtype=torch.float64, max_sifts=S, use_graph=False, device=dev)
    out = plan.forward([torch.tensor(x, device=dev)], noise=w)[0].cpu().numpy()
    gx = plan.backward([torch.tensor(G, device=dev)])[0].cpu().numpy()
    plan.close()
    ref_g, ref_imfs = orc.iceemdan_grad(x, w.cpu().numpy(), M, G, S=S)
    assert np.array_equal(out, ref_imfs)
    assert relerr(gx, ref_g) < 1e-9, relerr(gx, ref_g)


def test_backward_matches_gpu_finite_differences(icm, dev):
    """Directional derivative of L = sum_k w_k |IMF_k|^2 by central
    differences of the GPU forward (h = 1e-5)."""
    import torch
    N, I, M, S = 2048, 8, 5, 10
    x = _case(N, seed=4)
    rng = np.random.default_rng(5)
    wk = torch.tensor(rng.uniform(0, 1, M), device=dev)
    print("case: backward fd N=2048", flush=True)
    w = icm.noise(I, N, seed=3, dtype=torch.float64, device=dev)
    plan = icm.Plan([N], I, M, dtype=torch.float64, max_sifts=S, use_graph=False, device=dev)

    def loss(xx):
        im = plan.forward([torch.tensor(xx, device=dev)], noise=w)[0]
        return float((wk * (im ** 2).sum(1)).sum()), im

    _, im0 = loss(x)
    gx = plan.backward([2 * wk[:, None] * im0])[0].cpu().numpy()
    for _ in range(2):
        v = rng.standard_normal(N)
        v /= np.linalg.norm(v)
        hh = 1e-5
        fd = (loss(x + hh * v)[0] - loss(x - hh * v)[0]) / (2 * hh)
        assert abs(fd - gx @ v) <= 1e-6 * abs(fd), (fd, gx @ v)
    plan.close()


def test_backward_batched_equals_per_channel(icm, dev):
    """Uneven channels in one plan: each channel's gradient equals the one
    from a single-channel plan (same noise, shared per length)."""
    import torch
    sig = _sig()
    lens = [1500, 2200, 1500]
    I, M, S = 6, 4, 6
    xs = [torch.tensor(sig.eeg_like(n, seed=20 + c), device=dev) for c, n in enumerate(lens)]
    rng = np.random.default_rng(9)
    Gs = [torch.tensor(rng.standard_normal((M, n)), device=dev) for n in lens]
    print(f"case: backward batched lens={lens}", flush=True)
    plan = icm.Plan(lens, I, M, dtype=torch.float64, max_sifts=S, use_graph=False, device=dev)
    plan.forward(xs)
    gb = plan.backward(Gs)
    plan.close()
    for c, n in enumerate(lens):
        p1 = icm.Plan([n], I, M, dtype=torch.float64, max_sifts=S, use_graph=False, device=dev)
        p1.forward([xs[c]])
        g1 = p1.backward([Gs[c]])[0]
        p1.close()
        assert torch.equal(gb[c], g1), c


def test_autograd_function(icm, dev):
    import torch
    N, I, M, S = 1200, 5, 4, 8
    x = torch.tensor(_case(N, seed=8), device=dev, requires_grad=True)
    print("case: autograd", flush=True)
    plan = icm.Plan([N], I, M, dtype=torch.float64, max_sifts=S, use_graph=False, device=dev)
    imfs = icm.iceemdan_autograd(x, plan)
    loss = (imfs[0] ** 2).sum() + imfs[-1].sum()
    loss.backward()
    G = torch.zeros((M, N), dtype=torch.float64, device=dev)
    G[0] = 2 * imfs[0].detach()
    G[-1] = 1.0
    plan.forward([x.detach()])
    ref = plan.backward([G])[0]
    assert torch.allclose(x.grad, ref, rtol=1e-12, atol=1e-14)
    # a forward between forward and backward is detected
    y = torch.tensor(_case(N, seed=9), device=dev, requires_grad=True)
    out = icm.iceemdan_autograd(y, plan)
    plan.forward([x.detach()])
    with pytest.raises(RuntimeError):
        out.sum().backward()
    plan.close()


def test_backward_fp32_close_to_fp64(icm, orc, dev):
    import torch
    N, I, M, S = 2000, 8, 4, 8
    x = _case(N, seed=12)
    rng = np.random.default_rng(13)
    G = rng.standard_normal((M, N))
    print("case: backward fp32", flush=True)
    w = icm.noise(I, N, seed=1, dtype=torch.float64, device=dev)
    w32 = w.to(torch.float32)
    plan = icm.Plan([N], I, M, dtype=torch.float32, max_sifts=S, use_graph=False, device=dev)
    plan.forward([torch.tensor(x, dtype=torch.float32, device=dev)], noise=w32)
    gx = plan.backward([torch.tensor(G, dtype=torch.float32, device=dev)])[0]
    plan.close()
    gx = gx.cpu().numpy().astype(np.float64)
    ref, _ = orc.iceemdan_grad(x.astype(np.float32).astype(np.float64),
                               w32.cpu().numpy().astype(np.float64), M, G, S=S)
    print("fp32 backward rel err:", relerr(gx, ref))
    assert np.all(np.isfinite(gx))
    assert relerr(gx, ref) < 1e-2


def test_backward_state_errors(icm, dev):
    import torch
    N = 500
    x = torch.tensor(_case(N, seed=1), device=dev)
    G = torch.zeros((3, N), dtype=torch.float64, device=dev)
    plan = icm.Plan([N], 4, 3, dtype=torch.float64, max_sifts=5, use_graph=False, device=dev)
    with pytest.raises(icm.ICMError):
        plan.backward([G])          # no forward yet
    plan.close()
    plan = icm.Plan([N], 4, 3, dtype=torch.float64, stop_mode=icm.STOP_RILLING, max_sifts=20,
                    use_graph=False, device=dev)
    plan.forward([x])
    with pytest.raises(icm.ICMError):
        plan.backward([G])          # RILLING has no adjoint
    plan.close()


def test_keep_records_backward_equals_resift(icm, dev):
    """With keep_records the backward uses the forward's per-pass records
    instead of re-sifting; the gradient is the same bits.  After a resumed
    staged forward the stages run before the save are re-sifted."""
    import torch
    lens, I, M, S = [1800, 2600], 5, 5, 8
    xs = [torch.tensor(_case(n, seed=70 + n), device=dev) for n in lens]
    rng = np.random.default_rng(71)
    G = [torch.tensor(rng.standard_normal((M, n)), device=dev) for n in lens]
    print("case: keep_records", flush=True)
    ref_plan = icm.Plan(lens, I, M, dtype=torch.float64, max_sifts=S, use_graph=False, device=dev)
    ref_out = [o.clone() for o in ref_plan.forward(xs)]
    ref = ref_plan.backward(G)
    ref_plan.close()
    for graph in (False, True):
        kp = icm.Plan(lens, I, M, dtype=torch.float64, max_sifts=S, use_graph=graph, device=dev,
                      keep_records=True)
        for _ in range(2):   # the second forward replays the graph
            out = kp.forward(xs)
            for a, b in zip(out, ref_out):
                assert torch.equal(a, b)
            got = kp.backward(G)
            for a, b in zip(got, ref):
                assert torch.equal(a, b)
        # staged and resumed: stages 0..1 come from the state, 2.. are recorded
        kp.forward_begin(xs)
        kp.forward_stages(2)
        state = kp.save_state()
        kp.load_state(state)
        kp.forward_stages(100)
        kp.forward_finish()
        got = kp.backward(G)
        for a, b in zip(got, ref):
            assert torch.equal(a, b)
        kp.close()
