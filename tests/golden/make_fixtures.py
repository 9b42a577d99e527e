"""Writes oracle_fixtures.npz: outputs of the FP64 oracle on small fixed
cases (the reference ships no expected outputs).  The GPU parity tests
compare against the oracle run live; these fixtures pin the oracle itself
across edits (tests/test_oracle.py::test_golden_fixtures).

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def _eeg(n):
    return np.fromfile(os.path.join(HERE, "eegSampleDataCH4.f32"), dtype="<f4")[:n].astype(np.float64)


def _two_tone(N=1000):
    # the reference's synthetic example (sample_synthetic_signal.cu:48-55)
    n = np.arange(N)
    burst = np.where((n >= 500) & (n <= 748), np.sin(2 * np.pi * 0.255 * (n - 500)), 0.0)
    return np.sin(2 * np.pi * 0.065 * n) + burst


CASES = ["noise", "sift_eeg", "fixed_eeg", "rilling_two_tone", "grad_eeg"]


def run_case(orc, name):
    if name == "noise":
        return {"w": orc.noise(1234, 3, 37)}
    if name == "sift_eeg":
        h = _eeg(1500)
        hn, nmax, nmin, cnt, mark = orc.sift_pass(h)
        return {"h_new": hn, "counts": np.array([nmax, nmin, cnt, mark], np.int64)}
    if name == "fixed_eeg":
        x = _eeg(2000)
        w = orc.noise(7, 6, x.size)
        r = orc.iceemdan(x, w, 5, S_or_cap=10)
        return {"noise": w, "imfs": r["imfs"]}
    if name == "rilling_two_tone":
        x = _two_tone()
        w = orc.noise(1234, 8, x.size)
        r = orc.iceemdan(x, w, 3, stop_mode=orc.STOP_RILLING, S_or_cap=100, noise_cap=2000)
        return {"x": x, "noise": w, "imfs": r["imfs"], "stops": r["stops"]}
    if name == "grad_eeg":
        x = _eeg(800)
        w = orc.noise(3, 4, x.size)
        G = np.cos(np.arange(4 * x.size).reshape(4, x.size) / 17.0)
        gy, _ = orc.iceemdan_grad(x, w, 4, G, S=6)
        return {"noise": w, "grad": gy}
    raise KeyError(name)


def main():
    sys.path.insert(0, ROOT)
    from oracle import oracle as orc
    data = {}
    for name in CASES:
        for k, v in run_case(orc, name).items():
            data[f"{name}/{k}"] = v
    path = os.path.join(HERE, "oracle_fixtures.npz")
    np.savez_compressed(path, **data)
    print(path, sum(v.nbytes for v in data.values()), "bytes")


if __name__ == "__main__":
    main()
