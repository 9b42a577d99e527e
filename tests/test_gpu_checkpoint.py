"""Staged forward and stage-boundary checkpoint / resume: the result of a
resumed decomposition equals the uninterrupted one bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sig():
    import importlib
    return importlib.import_module("ceemdan-gpu_amd.signals")


def _inputs(dev, lens, seed=30):
    import torch
    return [torch.tensor(_sig().eeg_like(n, seed=seed + c), dtype=torch.float64, device=dev)
            for c, n in enumerate(lens)]


def test_staged_forward_equals_forward(icm, dev):
    import torch
    lens, I, M = [3000, 2100], 6, 6
    xs = _inputs(dev, lens)
    print("case: staged forward", flush=True)
    plan = icm.Plan(lens, I, M, dtype=torch.float64, max_sifts=10, use_graph=True, device=dev)
    ref = [o.clone() for o in plan.forward(xs)]
    assert plan.forward_begin(xs) == 0
    assert plan.forward_stages(2) == 2
    assert plan.forward_stages(100) == M - 1
    got = plan.forward_finish()
    for a, b in zip(got, ref):
        assert torch.equal(a, b)
    # the graph-replay forward still works after the staged one
    again = plan.forward(xs)
    for a, b in zip(again, ref):
        assert torch.equal(a, b)
    plan.close()


@pytest.mark.parametrize("stop_mode", [0, 1])
def test_resume_in_a_new_plan(icm, dev, stop_mode):
    import torch
    lens, I, M = [2500, 1800], 5, 5
    xs = _inputs(dev, lens, seed=40)
    kw = dict(dtype=torch.float64, stop_mode=stop_mode, max_sifts=10 if stop_mode == 0 else 40,
              noise_max_sifts=10 if stop_mode == 0 else 200, use_graph=False, device=dev)
    print(f"case: resume stop_mode={stop_mode}", flush=True)
    plan = icm.Plan(lens, I, M, **kw)
    ref = [o.clone() for o in plan.forward(xs)]
    plan.forward_begin(xs)
    plan.forward_stages(2)
    state = plan.save_state()
    plan.close()

    fresh = icm.Plan(lens, I, M, **kw)
    assert fresh.load_state(state) == 2
    fresh.forward_stages(100)
    got = fresh.forward_finish()
    for a, b in zip(got, ref):
        assert torch.equal(a, b)
    if stop_mode == 0:
        # the adjoint after a resumed forward equals the one after a full forward
        rng = np.random.default_rng(5)
        G = [torch.tensor(rng.standard_normal((M, n)), device=dev) for n in lens]
        g_resumed = fresh.backward(G)
        full = icm.Plan(lens, I, M, **kw)
        full.forward(xs)
        g_full = full.backward(G)
        full.close()
        for a, b in zip(g_resumed, g_full):
            assert torch.equal(a, b)
    fresh.close()


def test_resume_with_external_noise(icm, dev):
    import torch
    N, I, M = 2000, 4, 4
    x = _inputs(dev, [N], seed=50)
    w = icm.noise(I, N, seed=9, dtype=torch.float64, device=dev)
    print("case: resume external noise", flush=True)
    plan = icm.Plan([N], I, M, dtype=torch.float64, max_sifts=8, use_graph=False, device=dev)
    ref = plan.forward(x, noise=w)[0].clone()
    plan.forward_begin(x, noise=w)
    plan.forward_stages(1)
    state = plan.save_state()
    fresh = icm.Plan([N], I, M, dtype=torch.float64, max_sifts=8, use_graph=False, device=dev)
    with pytest.raises(icm.ICMError):
        fresh.load_state(state)          # the saved run was given its noise
    assert fresh.load_state(state, noise=w) == 1
    fresh.forward_stages(10)
    assert torch.equal(fresh.forward_finish()[0], ref)
    fresh.close()
    plan.close()


def test_state_rejects_other_configurations(icm, dev):
    import torch
    N = 1500
    x = _inputs(dev, [N], seed=60)
    plan = icm.Plan([N], 4, 4, dtype=torch.float64, max_sifts=8, use_graph=False, device=dev)
    with pytest.raises(icm.ICMError):
        plan.save_state()                # no staged forward in progress
    with pytest.raises(icm.ICMError):
        plan.forward_stages(1)
    plan.forward_begin(x)
    state = plan.save_state()
    other = icm.Plan([N], 5, 4, dtype=torch.float64, max_sifts=8, use_graph=False, device=dev)
    with pytest.raises(icm.ICMError):
        other.load_state(state)          # different number of realizations
    bad = state.copy()
    bad[0] ^= 0xFF
    with pytest.raises(icm.ICMError):
        plan.load_state(bad)             # not a state
    with pytest.raises(icm.ICMError):
        plan.forward_finish()            # stages not done
    other.close()
    plan.close()
