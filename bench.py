"""Benchmark of the ICEEMDAN hot path on one MI355X.

Default workload (the batched headline): 64 EEG-like channels with lengths
drawn uniformly from 8192..262144 samples (seed 3), each 1/f pink noise at
30 uV RMS plus 10 Hz alpha bursts (seeds 1000 + c), I = 200 noise
realizations per channel, eps0 = 0.2, M = 14 rows (13 sifted stages),
exactly 10 sifts per stage, float32, all channels and realizations in one
plan.

A "step" is one full forward: white-noise generation, the noise-mode stages
and the CEEMDAN stages, with the inputs already resident in HBM.
value = CEEMDAN sample-sifts of the step (sum over channels, stages and
realizations of N_c * 10) / step wall time.

Also in the one JSON line (rank 0):
  batched_vs_per_signal  the same 64 channels, one plan per channel run back
                         to back (the plain per-signal path), same settings
  roofline               the sift-pass kernel, HIP events on its stream over
                         the timed steps (the two phases run side by side,
                         so each launch shares the GPU with the other
                         phase's kernels); traffic from the committed
                         rocprofv3 FETCH_SIZE / WRITE_SIZE summary (its file
                         is named)
  roofline_alone         the same kernel with the GPU to itself: one extra
                         untimed forward with the phases one after the other
  fp32_vs_fp64           per-mode error of channel 0 against the FP64
                         decomposition (which equals the FP64 oracle bit for
                         bit)
  c2                     one EEG-like channel N = 244032, I = 500, M = 16,
                         float64 (the per-signal FP64 case)
  c4                     forward + adjoint, 32 x 16384, I = 100, M = 12,
                         float64: backward / forward time
  c5_share               (--c5) one GPU's share of the 128-channel FP32 job,
                         forward + adjoint
  cpu_baseline           the FP64 OpenMP oracle on a bounded sample of the
                         workload, best of 3, on this host's cores
"""
import argparse
import importlib
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ensemble sample-sifts/s batched vs per-signal (1 GPU); backward/forward time"
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    "c3": dict(I=200, M=14, S=10, dtype="f32",
               desc="64 EEG-like channels, N ~ U[8192, 262144] (seed 3), I=200, eps0=0.2, M=14, "
                    "10 sifts per stage, float32, one plan"),
    "c2": dict(I=500, M=16, S=10, dtype="f64",
               desc="one EEG-like channel N=244032 fs=500 Hz (pink 30 uV + 10 Hz alpha, seed 2), "
                    "I=500, eps0=0.2, M=16, 10 sifts per stage, float64"),
    "c1": dict(I=50, M=10, S=10, dtype="f64",
               desc="three-tone N=4096 fs=1 kHz + N(0,0.1^2) (seed 0), I=50, M=10, 10 sifts, float64"),
    "c4": dict(I=100, M=12, S=10, dtype="f64",
               desc="32 EEG-like channels x N=16384 (seed 4), I=100, M=12, 10 sifts, float64; "
                    "forward + adjoint of L = sum_k w_k |IMF_k|^2, w_k ~ U(0,1) (seed 5)"),
    # the 128-channel job shards channels over GPUs with no communication;
    # one GPU's share of an 8-way split is 16 channels
    "c5": dict(I=500, M=16, S=10, dtype="f32", records=True,
               desc="one GPU's share of the 128-channel job: channels 0-15 of 128 EEG-like x "
                    "N=244032 (+50 Hz 5 uV, seeds 6000+c), I=500, M=16, 10 sifts, float32; "
                    "forward + adjoint of L = sum_k w_k |IMF_k|^2"),
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


def traffic_file(workload, dtype):
    return os.path.join("profiles", f"k_sift_{workload}_fp{64 if dtype == 'f64' else 32}_traffic.md")


def measured_traffic(workload, dtype):
    """HBM bytes per sift launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (None when it was not profiled)."""
    path = os.path.join(ROOT, traffic_file(workload, dtype))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def _torch_dtype(torch, code):
    return torch.float64 if code == "f64" else torch.float32


def tune_forwards(plan, xd, out):
    """The plan's autotuning forwards (untimed, like plan creation): the sift
    candidates are timed in rotation over the passes of these forwards.  A
    plan whose shape this process has tuned already starts tuned."""
    t0 = time.perf_counter()
    n = 0
    while plan.tuning_left() > 0 and n < 8:
        plan.forward(xd, out=out)
        n += 1
    return {"forwards": n, "ms": 1e3 * (time.perf_counter() - t0)}


def tuning_of(icm, plan, tune_ms):
    """The autotuner's choice per stage: (tiles per segment, small tables)."""
    desc = {}
    for ph, name in ((icm.PHASE_NOISE, "noise_modes"), (icm.PHASE_SIGNAL, "ceemdan")):
        desc[name] = [list(icm.tuning_candidate(int(c))) for c in plan.tuning(ph)]
    desc["tuning_forwards"] = tune_ms
    return desc


def forward_workload(icm, torch, dev, name, xs, steps, warmup, graph=False):
    """Times `steps` forwards of one plan over all channels (after `warmup`
    untimed ones); the sift kernel's launches are timed with HIP events."""
    import numpy as np
    W = WORKLOADS[name]
    dtype = _torch_dtype(torch, W["dtype"])
    lengths = [x.size for x in xs]
    I, M, S = W["I"], W["M"], W["S"]
    plan = icm.Plan(lengths, I, M, dtype=dtype, max_sifts=S, noise_strength=0.2, seed=1234,
                    use_graph=graph, device=dev)
    xd = [torch.tensor(x, dtype=dtype, device=dev) for x in xs]
    out = [torch.empty((M, n), dtype=dtype, device=dev) for n in lengths]
    tune_ms = tune_forwards(plan, xd, out)
    plan.set_profiling(not graph)
    for _ in range(warmup):
        plan.forward(xd, out=out)
    torch.cuda.synchronize()
    sift_ms = sift_n = sift_bytes = 0.0
    totals = []
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.forward(xd, out=out)
        totals.append(plan.timing())
        if not graph:
            ks = plan.kernel_stats()
            sift_ms += ks["sift"][0]
            sift_n += ks["sift"][1]
            sift_bytes += ks["sift"][2]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sample_sifts = (M - 1) * I * sum(lengths) * S
    res = {
        "plan": plan, "xd": xd, "out": out, "tuning": tuning_of(icm, plan, tune_ms),
        "value": sample_sifts / (wall / steps),
        "ms_per_step": 1e3 * wall / steps,
        "sample_sifts_per_step": sample_sifts,
        "phases_ms": {"noise_modes": float(np.mean([t["noise_ms"] for t in totals])),
                      "ceemdan": float(np.mean([t["signal_ms"] for t in totals])),
                      "total": float(np.mean([t["total_ms"] for t in totals]))},
        "output_finite": all(bool(torch.isfinite(o).all()) for o in out),
    }
    res["ceemdan_phase_rate"] = sample_sifts / (1e-3 * res["phases_ms"]["ceemdan"])
    if not graph and sift_n:
        # the same sift launches with the kernel alone on the GPU: one more
        # (untimed) forward with the two phases one after the other
        plan.set_overlap(False)
        plan.forward(xd, out=out)
        torch.cuda.synchronize()
        ks = plan.kernel_stats()
        plan.set_overlap(True)
        alone_ms = ks["sift"][0] / ks["sift"][1]
        alone_bpl = ks["sift"][2] / ks["sift"][1]
        alone = alone_bpl / (alone_ms * 1e-3) / 1e9
        res["roofline_alone"] = {"achieved": alone, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                 "frac": alone / PEAK_HBM_GBS, "avg_launch_ms": alone_ms,
                                 "launches": int(ks["sift"][1]),
                                 "note": "one extra forward, phases one after the other"}
        avg_ms = sift_ms / sift_n
        bpl = sift_bytes / sift_n
        achieved = bpl / (avg_ms * 1e-3) / 1e9
        res["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                           "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                           "traffic": measured_traffic(name, W["dtype"]),
                           "traffic_source": traffic_file(name, W["dtype"])
                           if measured_traffic(name, W["dtype"]) else None,
                           "kernel": "k_sift_w (+ k_sift_list on small-table passes)",
                           "avg_launch_ms": avg_ms, "alg_bytes_per_launch": bpl,
                           "launches": int(sift_n),
                           "note": "timed steps run the noise-mode and CEEMDAN phases side by "
                                   "side, so each sift launch shares the CUs with the other "
                                   "phase's kernels; roofline_alone is the same kernel with "
                                   "the GPU to itself"}
    return res


def graph_replay_ms(icm, torch, dev, name, xs, steps):
    """ms per forward of the hipGraph replay (no per-kernel events)."""
    W = WORKLOADS[name]
    dtype = _torch_dtype(torch, W["dtype"])
    lengths = [x.size for x in xs]
    plan = icm.Plan(lengths, W["I"], W["M"], dtype=dtype, max_sifts=W["S"], use_graph=True,
                    device=dev)
    xd = [torch.tensor(x, dtype=dtype, device=dev) for x in xs]
    out = [torch.empty((W["M"], n), dtype=dtype, device=dev) for n in lengths]
    tune_forwards(plan, xd, out)
    plan.forward(xd, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.forward(xd, out=out)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    plan.close()
    return ms


def per_signal(icm, torch, dev, name, xs, batched_value):
    """The plain per-signal path on the same channels: one plan per channel
    (hipGraph replay, same settings, autotuned like the batched plan), each
    timed on the forward after the one that captures its graph, run back to
    back.  Plan creation and tuning are outside the timed region, as the
    batched plan's are."""
    W = WORKLOADS[name]
    dtype = _torch_dtype(torch, W["dtype"])
    I, M, S = W["I"], W["M"], W["S"]
    total = 0.0
    for x in xs:
        n = x.size
        p = icm.Plan([n], I, M, dtype=dtype, max_sifts=S, use_graph=True, device=dev)
        xd = [torch.tensor(x, dtype=dtype, device=dev)]
        out = [torch.empty((M, n), dtype=dtype, device=dev)]
        tune_forwards(p, xd, out)
        p.forward(xd, out=out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p.forward(xd, out=out)
        torch.cuda.synchronize()
        total += time.perf_counter() - t0
        p.close()
    sifts = (M - 1) * I * sum(x.size for x in xs) * S
    ps = sifts / total
    return {"workload": f"the same {len(xs)} channels, one plan per channel, hipGraph replay",
            "batched": batched_value, "per_signal": ps, "ratio": batched_value / ps,
            "per_signal_ms": 1e3 * total}


def fp32_error(icm, torch, dev, name, xs, out0):
    """Per-mode error of the FP32 decomposition of channel 0 against the
    FP64 decomposition of the same float32 input and the same float32 noise
    (the FP64 path equals the FP64 oracle bit for bit), same settings."""
    W = WORKLOADS[name]
    I, M, S = W["I"], W["M"], W["S"]
    n0 = xs[0].size
    p64 = icm.Plan([n0], I, M, dtype=torch.float64, max_sifts=S, use_graph=False, device=dev)
    x32 = torch.tensor(xs[0], dtype=torch.float32, device=dev)
    w32 = icm.noise(I, n0, seed=1234, dtype=torch.float32, device=dev)
    ref = p64.forward([x32.to(torch.float64)], noise=w32.to(torch.float64))[0]
    p64.close()
    got = out0.to(torch.float64)
    # null where the FP64 row is zero (a stage with too few extrema
    # subtracts nothing, and its residue row is exactly 0)
    rel = [float((got[k] - ref[k]).abs().max() / ref[k].abs().max())
           if float(ref[k].abs().max()) > 0 else None for k in range(M)]
    rms = [float((got[k] - ref[k]).norm() / ref[k].norm())
           if float(ref[k].norm()) > 0 else None for k in range(M)]
    return {"channel": 0, "samples": n0, "reference": "FP64 decomposition of the float32 input",
            "per_mode_max_rel": rel, "per_mode_rms_rel": rms}


def backward_workload(icm, torch, dev, name, xs, steps):
    """Forward + adjoint of L = sum_k w_k |IMF_k|^2 (w_k ~ U(0,1), seed 5)
    with a plan that keeps every pass's knot records."""
    import numpy as np
    W = WORKLOADS[name]
    dtype = _torch_dtype(torch, W["dtype"])
    lengths = [x.size for x in xs]
    I, M, S = W["I"], W["M"], W["S"]
    records = W.get("records", True)
    plan = icm.Plan(lengths, I, M, dtype=dtype, max_sifts=S, noise_strength=0.2, seed=1234,
                    use_graph=False, device=dev, keep_records=records)
    xd = [torch.tensor(x, dtype=dtype, device=dev) for x in xs]
    out = [torch.empty((M, n), dtype=dtype, device=dev) for n in lengths]
    wk = torch.tensor(np.random.default_rng(5).uniform(0, 1, M), dtype=dtype, device=dev)
    tune_forwards(plan, xd, out)
    plan.forward(xd, out=out)
    grads = [2 * wk[:, None] * o for o in out]
    gx = plan.backward(grads)
    torch.cuda.synchronize()
    fwd = bwd = 0.0
    for _ in range(steps):
        t1 = time.perf_counter()
        plan.forward(xd, out=out)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        gx = plan.backward(grads)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        fwd += t2 - t1
        bwd += t3 - t2
    plan.close()
    sifts = (M - 1) * I * sum(lengths) * S
    fwd_ms, bwd_ms = 1e3 * fwd / steps, 1e3 * bwd / steps
    return {"workload": W["desc"], "forward_ms": fwd_ms, "backward_ms": bwd_ms,
            "backward_forward_ratio": bwd_ms / fwd_ms, "keep_records": records,
            "forward_sample_sifts_per_s": sifts / (1e-3 * fwd_ms),
            "grad_finite": all(bool(torch.isfinite(g).all()) for g in gx)}


def host_cpu():
    """(physical cores, model name) of this host from lscpu."""
    cores, model = os.cpu_count() or 1, "unknown"
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True).stdout
        pairs = {l for l in out.splitlines() if l and not l.startswith("#")}
        if pairs:
            cores = len(pairs)
        for l in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if l.startswith("Model name:"):
                model = l.split(":", 1)[1].strip()
    except OSError:
        pass
    return cores, model


def cpu_baseline(x, I_full, M, S, sample_I, what):
    """FP64 oracle (OpenMP over realizations) on sample_I of the
    realizations of one channel, best of 3 after one warm-up run."""
    from oracle import oracle as orc

    phys, model = host_cpu()
    threads = min(int(os.environ.get("OMP_NUM_THREADS", phys)), phys)
    w = orc.noise(1234, sample_I, x.size)
    orc.iceemdan(x[: min(x.size, 4096)], w[:, : min(x.size, 4096)], M, S_or_cap=S,
                 nthreads=threads)
    best = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        orc.iceemdan(x, w, M, S_or_cap=S, nthreads=threads)
        best = min(best, time.perf_counter() - t0)
    sifts = (M - 1) * sample_I * x.size * S
    return {"value": sifts / best, "unit": "sample-sifts/s", "cores": threads, "kind": "port",
            "physical_cores": phys, "cpu_model": model,
            "sample": f"FP64 oracle, {what} N={x.size}, {sample_I} of {I_full} realizations, "
                      f"M={M}, {S} sifts (noise-mode + CEEMDAN stages, CEEMDAN sample-sifts "
                      f"counted), best of 3: {best:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample", type=int, default=48, help="realizations timed on the CPU")
    ap.add_argument("--no-per-signal", action="store_true",
                    help="skip the batched vs per-signal comparison")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the C2 / C4 sub-measurements and the graph replay")
    ap.add_argument("--c5", action="store_true", help="also time the C5 share forward + adjoint")
    ap.add_argument("--forward-only", action="store_true",
                    help="only the timed forwards of the workload (for profiler runs)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the forward + adjoint timing of the workload, one JSON line "
                         "(for profiler runs)")
    args = ap.parse_args()
    if args.forward_only:
        args.no_cpu = args.no_per_signal = args.no_extra = True

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(channels would shard across GPUs with no communication).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    icm = importlib.import_module("ceemdan-gpu_amd")
    sig = importlib.import_module("ceemdan-gpu_amd.signals")
    dev = torch.device("cuda", 0)
    name = args.workload
    W = WORKLOADS[name]
    xs = signals_for(name, sig)
    lengths = [x.size for x in xs]
    if args.adjoint_only:
        print(json.dumps({"workload": name, **backward_workload(icm, torch, dev, name, xs,
                                                                args.steps)}))
        return

    main_res = forward_workload(icm, torch, dev, name, xs, args.steps, args.warmup)
    main_res["plan"].close()
    line = {
        "metric": METRIC,
        "value": main_res["value"],
        "unit": "sample-sifts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": W["dtype"],
        "data": "synthetic",
        "config": {"workload": W["desc"], "channels": len(lengths), "samples": sum(lengths),
                   "realizations": W["I"], "rows": W["M"], "sifts_per_stage": W["S"],
                   "eps0": 0.2, "sample_sifts_per_step": main_res["sample_sifts_per_step"]},
        "phases_ms": main_res["phases_ms"],
        "ceemdan_phase_rate": main_res["ceemdan_phase_rate"],
        "output_finite": main_res["output_finite"],
    }
    if "roofline" in main_res:
        line["roofline"] = main_res["roofline"]
        line["roofline_alone"] = main_res["roofline_alone"]
    line["autotune"] = main_res["tuning"]
    if W["dtype"] == "f32" and not args.forward_only:
        line["fp32_vs_fp64"] = fp32_error(icm, torch, dev, name, xs, main_res["out"][0])
    del main_res
    torch.cuda.empty_cache()
    if not args.no_extra:
        line["graph_replay_ms_per_step"] = graph_replay_ms(icm, torch, dev, name, xs, 2)
    if not args.no_per_signal and len(xs) > 1:
        line["batched_vs_per_signal"] = per_signal(icm, torch, dev, name, xs, line["value"])
    if not args.no_extra:
        if name != "c2":
            c2 = forward_workload(icm, torch, dev, "c2", signals_for("c2", sig), 3, 1)
            c2["plan"].close()
            line["c2"] = {"workload": WORKLOADS["c2"]["desc"], "value": c2["value"],
                          "ms_per_step": c2["ms_per_step"],
                          "roofline_frac": c2.get("roofline", {}).get("frac"),
                          "sift_avg_launch_ms": c2.get("roofline", {}).get("avg_launch_ms"),
                          "roofline_frac_alone": c2.get("roofline_alone", {}).get("frac"),
                          "sift_avg_launch_ms_alone": c2.get("roofline_alone", {}).get("avg_launch_ms")}
            del c2
            torch.cuda.empty_cache()
        line["c4"] = backward_workload(icm, torch, dev, "c4", signals_for("c4", sig), 3)
        torch.cuda.empty_cache()
    if args.c5:
        line["c5_share"] = backward_workload(icm, torch, dev, "c5", signals_for("c5", sig), 1)
        torch.cuda.empty_cache()
    if not args.no_cpu:
        sample_I = max(1, min(args.cpu_sample, W["I"]))
        line["cpu_baseline"] = cpu_baseline(np.asarray(xs[0], np.float64), W["I"], W["M"], W["S"],
                                            sample_I, f"channel 0 of the {name} workload")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
