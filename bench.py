"""Benchmark of the ICEEMDAN hot path on one MI355X.

Default workload (one EEG-like channel, the plain per-signal FP64 case):
N = 244032 samples at 500 Hz (1/f pink noise at 30 uV RMS + 10 Hz alpha
bursts, seed 2), I = 500 noise realizations, eps0 = 0.2, M = 16 rows
(15 sifted stages), exactly 10 sifts per stage, float64.

A "step" is one full forward: white-noise generation, the noise-mode stages
E_1..E_15 (also 10 sifts each) and the 15 CEEMDAN stages, with the input
already resident in HBM.  value = CEEMDAN sample-sifts of the step
(sum over stages and realizations of N * 10 = 1.83e10) / step wall time.

Prints one JSON line (rank 0).  `roofline` is the sift-pass kernel measured
with HIP events on its stream over the timed steps; `cpu_baseline` is the
FP64 OpenMP oracle on a bounded sample of the same workload (a subset of the
realizations) on this host's cores.
"""
import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ensemble sample-sifts/s batched vs per-signal (1 GPU); backward/forward time"
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    # name: (channels lengths builder, I, M, S, dtype, description)
    "c2": dict(I=500, M=16, S=10, dtype="f64",
               desc="one EEG-like channel N=244032 fs=500 Hz (pink 30 uV + 10 Hz alpha, seed 2), "
                    "I=500, eps0=0.2, M=16, 10 sifts per stage, float64"),
    "c1": dict(I=50, M=10, S=10, dtype="f64",
               desc="three-tone N=4096 fs=1 kHz + N(0,0.1^2) (seed 0), I=50, M=10, 10 sifts, float64"),
    "c3": dict(I=200, M=14, S=10, dtype="f32",
               desc="64 EEG-like channels, N ~ U[8192, 262144] (seed 3), I=200, M=14, 10 sifts, float32"),
    "c4": dict(I=100, M=12, S=10, dtype="f64",
               desc="32 EEG-like channels x N=16384 (seed 4), I=100, M=12, 10 sifts, float64; "
                    "forward + adjoint of L = sum_k w_k |IMF_k|^2, w_k ~ U(0,1) (seed 5)"),
    # the 128-channel job shards channels over 8 GPUs with no communication;
    # one GPU's share is 16 channels (all 128 in one plan exceed the 2^32
    # work-items of one launch).  Per-pass knot records would take 1.26 TB,
    # so the backward re-sifts each stage.
    "c5": dict(I=500, M=16, S=10, dtype="f32", records=True,
               desc="one GPU's share of the 128-channel job: channels 0-15 of 128 EEG-like x "
                    "N=244032 (+50 Hz 5 uV, seeds 6000+c), I=500, M=16, 10 sifts, float32; "
                    "forward + adjoint of L = sum_k w_k |IMF_k|^2 (backward re-sifts)"),
}


def signals_for(name, sig):
    if name == "c2":
        return [sig.eeg_like(244032, seed=2)]
    if name == "c1":
        return [sig.three_tone(4096, seed=0)]
    if name == "c3":
        lens = sig.batch_lengths(64, 8192, 262144, seed=3)
        return [sig.eeg_like(n, seed=1000 + c) for c, n in enumerate(lens)]
    if name == "c4":
        return [sig.eeg_like(16384, seed=4000 + c) for c in range(32)]
    if name == "c5":
        return [sig.eeg_like(244032, seed=6000 + c, line_uv=5.0) for c in range(16)]
    raise ValueError(name)


def measured_traffic(workload, dtype):
    """HBM bytes per k_sift launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (None when it was not profiled)."""
    path = os.path.join(ROOT, "profiles", f"k_sift_{workload}_fp{64 if dtype == 'f64' else 32}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def batched_vs_per_signal(icm, sig, torch, dev, channels=8, N=16384, I=100, M=12, S=10):
    """Sample-sifts/s of one plan over `channels` channels against one plan
    per channel run back to back (same FIXED settings, float64)."""
    xs = [torch.tensor(sig.eeg_like(N, seed=4000 + c), device=dev) for c in range(channels)]
    sifts = (M - 1) * I * N * S * channels
    plan = icm.Plan([N] * channels, I, M, dtype=torch.float64, max_sifts=S, use_graph=True,
                    device=dev)
    plan.forward(xs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.forward(xs)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    plan.close()
    single = icm.Plan([N], I, M, dtype=torch.float64, max_sifts=S, use_graph=True, device=dev)
    single.forward([xs[0]])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for x in xs:
        single.forward([x])
    torch.cuda.synchronize()
    ts = time.perf_counter() - t0
    single.close()
    return {"workload": f"{channels} EEG-like channels x N={N}, I={I}, M={M}, {S} sifts, float64",
            "batched": sifts / tb, "per_signal": sifts / ts, "ratio": ts / tb}


def cpu_baseline(x, I_full, M, S, sample_I, threads):
    """FP64 oracle (OpenMP over realizations) on sample_I of the realizations."""
    import numpy as np
    from oracle import oracle as orc

    w = orc.noise(1234, sample_I, x.size)
    t0 = time.perf_counter()
    orc.iceemdan(x, w, M, S_or_cap=S, nthreads=threads)
    dt = time.perf_counter() - t0
    sifts = (M - 1) * sample_I * x.size * S
    return {"value": sifts / dt, "unit": "sample-sifts/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle, one channel N={x.size}, {sample_I} of {I_full} realizations, "
                      f"M={M}, {S} sifts (noise-mode + CEEMDAN stages), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample", type=int, default=24, help="realizations timed on the CPU")
    ap.add_argument("--graph", action="store_true", help="time the hipGraph replay (no kernel events)")
    ap.add_argument("--no-backward", action="store_true", help="skip the adjoint timing")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched vs per-signal comparison")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(channels would shard across GPUs with no communication).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    icm = importlib.import_module("ceemdan-gpu_amd")
    sig = importlib.import_module("ceemdan-gpu_amd.signals")
    W = WORKLOADS[args.workload]
    dtype = torch.float64 if W["dtype"] == "f64" else torch.float32
    dev = torch.device("cuda", 0)
    xs = signals_for(args.workload, sig)
    lengths = [x.size for x in xs]
    I, M, S = W["I"], W["M"], W["S"]
    plan = icm.Plan(lengths, I, M, dtype=dtype, max_sifts=S, noise_strength=0.2, seed=1234,
                    use_graph=args.graph, device=dev)
    plan.set_profiling(not args.graph)
    xd = [torch.tensor(x, dtype=dtype, device=dev) for x in xs]
    out = [torch.empty((M, n), dtype=dtype, device=dev) for n in lengths]

    for _ in range(args.warmup):
        plan.forward(xd, out=out)
    torch.cuda.synchronize()

    sift_ms = sift_n = sift_bytes = 0.0
    totals = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.forward(xd, out=out)
        tm = plan.timing()
        totals.append(tm)
        if not args.graph:
            ks = plan.kernel_stats()
            sift_ms += ks["sift"][0]
            sift_n += ks["sift"][1]
            sift_bytes += ks["sift"][2]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    sample_sifts = (M - 1) * I * sum(lengths) * S
    ms_per_step = 1e3 * wall / args.steps
    value = sample_sifts / (wall / args.steps)
    finite = all(bool(torch.isfinite(o).all()) for o in out)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "sample-sifts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": W["dtype"],
        "data": "synthetic",
        "config": {"workload": W["desc"], "channels": len(lengths), "samples": sum(lengths),
                   "realizations": I, "rows": M, "sifts_per_stage": S, "eps0": 0.2,
                   "sample_sifts_per_step": sample_sifts},
        "phases_ms": {"noise_modes": float(np.mean([t["noise_ms"] for t in totals])),
                      "ceemdan": float(np.mean([t["signal_ms"] for t in totals])),
                      "total": float(np.mean([t["total_ms"] for t in totals]))},
        "ceemdan_phase_rate": sample_sifts / (1e-3 * float(np.mean([t["signal_ms"] for t in totals]))),
        "output_finite": finite,
    }
    if not args.graph and sift_n:
        avg_ms = sift_ms / sift_n
        bytes_per_launch = sift_bytes / sift_n
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        line["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                            "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                            "traffic": measured_traffic(args.workload, W["dtype"]),
                            "kernel": "k_sift", "avg_launch_ms": avg_ms,
                            "alg_bytes_per_launch": bytes_per_launch,
                            "launches": int(sift_n)}
    if not args.no_backward:
        # adjoint of L = sum_k w_k |IMF_k|^2, w_k ~ U(0,1) (seed 5), with a plan
        # whose forward keeps every pass's knot structure (keep_records), so
        # the backward does not re-sift (unless the records do not fit); its
        # forward is timed as well
        records = W.get("records", True)
        plan.close()
        plan = icm.Plan(lengths, I, M, dtype=dtype, max_sifts=S, noise_strength=0.2, seed=1234,
                        use_graph=False, device=dev, keep_records=records)
        wk = torch.tensor(np.random.default_rng(5).uniform(0, 1, M), dtype=dtype, device=dev)
        plan.forward(xd, out=out)
        grads = [2 * wk[:, None] * o for o in out]
        gx = plan.backward(grads)
        torch.cuda.synchronize()
        fwd_r = bwd = 0.0
        for _ in range(args.steps):
            t1 = time.perf_counter()
            plan.forward(xd, out=out)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            gx = plan.backward(grads)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            fwd_r += t2 - t1
            bwd += t3 - t2
        fwd_ms = 1e3 * fwd_r / args.steps
        bwd_ms = 1e3 * bwd / args.steps
        line["backward"] = {"ms_per_call": bwd_ms, "forward_ms": fwd_ms, "keep_records": records,
                            "backward_forward_ratio": bwd_ms / fwd_ms,
                            "backward_vs_plain_forward": bwd_ms / ms_per_step,
                            "loss": "sum_k w_k |IMF_k|^2, w_k ~ U(0,1) seed 5",
                            "grad_finite": all(bool(torch.isfinite(g).all()) for g in gx)}
    if W["dtype"] == "f32":
        # per-mode error of the FP32 decomposition against FP64 (channel 0,
        # same settings; the FP64 GPU path equals the FP64 oracle bit for bit)
        n0 = lengths[0]
        p64 = icm.Plan([n0], I, M, dtype=torch.float64, max_sifts=S, noise_strength=0.2, seed=1234,
                       use_graph=False, device=dev)
        ref = p64.forward([torch.tensor(xs[0], dtype=torch.float64, device=dev)])[0]
        p64.close()
        got = out[0].to(torch.float64)
        # (null where the FP64 row is zero: a stage with too few extrema
        # subtracts nothing, and its residue row is exactly 0)
        rel = [float((got[k] - ref[k]).abs().max() / ref[k].abs().max())
               if float(ref[k].abs().max()) > 0 else None for k in range(M)]
        rms = [float((got[k] - ref[k]).norm() / ref[k].norm())
               if float(ref[k].norm()) > 0 else None for k in range(M)]
        line["fp32_vs_fp64"] = {"channel": 0, "samples": n0, "per_mode_max_rel": rel,
                                "per_mode_rms_rel": rms}
    if not args.no_batched:
        line["batched_vs_per_signal"] = batched_vs_per_signal(icm, sig, torch, dev)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        sample_I = max(1, min(args.cpu_sample, I))
        line["cpu_baseline"] = cpu_baseline(np.asarray(xs[0], np.float64), I, M, S, sample_I,
                                            threads)
    plan.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
