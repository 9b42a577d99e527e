"""The sift autotuner, the work-queue sift and the grid-stride tile kernels
on the GPU.

Every tuning candidate (segment length x knot-table size) must compute the
same bits as the default geometry, per stage and mixed across stages, in
both precisions and both stop rules; the FP64 result is also the oracle's.
A tuning run must end with a choice per stage, keep the bits, and leave its
choice to later plans of the same shape."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sig():
    return importlib.import_module("ceemdan-gpu_amd.signals")


def _inputs(dtype, lens):
    import torch
    sig = _sig()
    return [torch.tensor(sig.eeg_like(n, seed=300 + c), dtype=dtype, device="cuda")
            for c, n in enumerate(lens)]


@pytest.mark.parametrize("dt,stop", [("f64", 0), ("f32", 0), ("f64", 1), ("f32", 1)])
def test_every_candidate_computes_the_same_bits(icm, dev, dt, stop):
    import torch
    dtype = torch.float64 if dt == "f64" else torch.float32
    lens, I, M, S = [3000, 70000, 4500, 70000], 4, 5, 6
    print(f"case: tuning candidates {dt} stop={stop} lens={lens}", flush=True)
    xs = _inputs(dtype, lens)
    plan = icm.Plan(lens, I, M, dtype=dtype, stop_mode=stop, max_sifts=S, use_graph=False,
                    device=dev, noise_max_sifts=S if stop == 0 else 50)
    plan.set_autotune(0)
    ref = [o.clone() for o in plan.forward(xs)]
    assert list(plan.tuning(icm.PHASE_SIGNAL)) == [-1] * (M - 1)
    for c in range(icm.NUM_TUNING_CANDIDATES):
        for ph in (icm.PHASE_NOISE, icm.PHASE_SIGNAL):
            plan.set_tuning(ph, [c] * (M - 1))
        out = plan.forward(xs)
        for a, b in zip(out, ref):
            assert torch.equal(a, b), c
    # a different candidate per stage
    rng = np.random.default_rng(1)
    for ph in (icm.PHASE_NOISE, icm.PHASE_SIGNAL):
        plan.set_tuning(ph, rng.integers(-1, icm.NUM_TUNING_CANDIDATES, M - 1))
    out = plan.forward(xs)
    plan.close()
    for a, b in zip(out, ref):
        assert torch.equal(a, b)


def test_candidates_match_the_oracle(icm, orc, dev):
    """FP64 under two extreme candidates (4-tile segments with small tables,
    32-tile segments with full tables) equals the oracle per channel."""
    import torch
    lens, I, M, S = [70000, 3000], 3, 4, 5
    sig = _sig()
    xs = [sig.eeg_like(n, seed=310 + c) for c, n in enumerate(lens)]
    print(f"case: tuning candidates vs oracle lens={lens}", flush=True)
    plan = icm.Plan(lens, I, M, dtype=torch.float64, max_sifts=S, use_graph=False, device=dev)
    for cand in (1, 6):
        for ph in (icm.PHASE_NOISE, icm.PHASE_SIGNAL):
            plan.set_tuning(ph, [cand] * (M - 1))
        out = plan.forward([torch.tensor(x, device=dev) for x in xs])
        for c, (n, x) in enumerate(zip(lens, xs)):
            w = icm.noise(I, n, seed=1234, dtype=torch.float64, device=dev).cpu().numpy()
            ref = orc.iceemdan(x, w, M, S_or_cap=S)["imfs"]
            assert np.array_equal(out[c].cpu().numpy(), ref), (cand, c)
    plan.close()


def test_autotune_run_and_table(icm, dev):
    """Two tuning forwards pick a candidate for every stage without changing
    the result; the graph replay after tuning equals them; a new plan of the
    same shape starts with the same choice."""
    import torch
    lens, I, M, S = [9000, 40000, 9000], 6, 5, 8
    print(f"case: autotune run lens={lens}", flush=True)
    xs = _inputs(torch.float32, lens)
    plan = icm.Plan(lens, I, M, dtype=torch.float32, max_sifts=S, use_graph=True, device=dev)
    plan.set_autotune(0)
    ref = [o.clone() for o in plan.forward(xs)]
    plan.set_autotune(2)
    for _ in range(2):
        out = plan.forward(xs)
        for a, b in zip(out, ref):
            assert torch.equal(a, b)
    tuned = {ph: plan.tuning(ph) for ph in (icm.PHASE_NOISE, icm.PHASE_SIGNAL)}
    for ph, t in tuned.items():
        assert np.all((t >= 0) & (t < icm.NUM_TUNING_CANDIDATES)), (ph, t)
    for _ in range(2):   # captured graph, replayed
        out = plan.forward(xs)
        torch.cuda.synchronize()
        for a, b in zip(out, ref):
            assert torch.equal(a, b)
    plan.close()
    again = icm.Plan(lens, I, M, dtype=torch.float32, max_sifts=S, use_graph=True, device=dev)
    for ph, t in tuned.items():
        assert np.array_equal(again.tuning(ph), t), ph
    out = again.forward(xs)
    again.close()
    for a, b in zip(out, ref):
        assert torch.equal(a, b)


@pytest.mark.parametrize("stop", [0, 1])
def test_queue_and_grid_stride_instances_match(icm, orc, dev, monkeypatch, stop):
    """The instances a plan uses only past a launch's 2^32 work-items (the
    work-queue sift over a resident grid, the grid-stride stage start and
    adjoint tile step), forced on a small plan with ICM_SIFT_QUEUE=1 and
    ICM_GRID_LOOP=1: the same forward bits as the default instances, equal
    to the oracle per channel (FP64), and in FIXED mode the same adjoint
    bits."""
    import torch
    lens, I, M, S = [70000, 3000], 3, 4, 5
    sig = _sig()
    xs = [sig.eeg_like(n, seed=320 + c) for c, n in enumerate(lens)]
    xd = [torch.tensor(x, device=dev) for x in xs]
    rng = np.random.default_rng(7)
    grads = [torch.tensor(rng.standard_normal((M, n)), device=dev) for n in lens]
    print(f"case: queue + grid-stride instances stop={stop} lens={lens}", flush=True)

    def run(forced):
        for k in ("ICM_SIFT_QUEUE", "ICM_GRID_LOOP"):
            if forced:
                monkeypatch.setenv(k, "1")
            else:
                monkeypatch.delenv(k, raising=False)
        plan = icm.Plan(lens, I, M, dtype=torch.float64, stop_mode=stop, max_sifts=S,
                        noise_max_sifts=S if stop == 0 else 50, use_graph=False, device=dev,
                        keep_records=stop == 0)
        out = [o.clone() for o in plan.forward(xd)]
        g = [t.clone() for t in plan.backward(grads)] if stop == 0 else []
        plan.close()
        return out, g

    ref_out, ref_g = run(False)
    out, g = run(True)
    monkeypatch.delenv("ICM_SIFT_QUEUE", raising=False)
    monkeypatch.delenv("ICM_GRID_LOOP", raising=False)
    for a, b in zip(out, ref_out):
        assert torch.equal(a, b)
    for a, b in zip(g, ref_g):
        assert torch.equal(a, b)
    for c, (n, x) in enumerate(zip(lens, xs)):
        w = icm.noise(I, n, seed=1234, dtype=torch.float64, device=dev).cpu().numpy()
        ref = orc.iceemdan(x, w, M, stop_mode=stop, S_or_cap=S,
                           noise_cap=S if stop == 0 else 50)["imfs"]
        assert np.array_equal(out[c].cpu().numpy(), ref), c
