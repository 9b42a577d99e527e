"""Synthetic test and benchmark signals (host side, numpy).

* two_tone(N): the reference's synthetic example (sample_synthetic_signal.cu:48-55):
  sin(2 pi 0.065 n), plus sin(2 pi 0.255 (n - 500)) on 500 <= n <= 748.
* three_tone(N, seed): sin(2 pi 5 t) + 0.5 sin(2 pi 20 t) + 0.25 sin(2 pi 60 t)
  + N(0, 0.1^2), t = n / 1 kHz.
* eeg_like(N, seed, fs=500): 1/f (pink) noise shaped in the FFT domain and
  scaled to 30 uV RMS, plus 10 Hz alpha bursts of 20 uV with Gaussian
  envelopes of 0.5 s FWHM at Poisson rate 0.5/s, optional line component.
* batch_lengths(C, lo, hi, seed): channel lengths drawn uniformly.
"""
import numpy as np


def two_tone(N=1000):
    n = np.arange(N)
    y = np.sin(2 * np.pi * 0.065 * n)
    burst = (n >= 500) & (n <= 748)
    y = y + np.where(burst, np.sin(2 * np.pi * 0.255 * (n - 500)), 0.0)
    return y


def three_tone(N=4096, seed=0, fs=1000.0):
    t = np.arange(N) / fs
    rng = np.random.default_rng(seed)
    return (np.sin(2 * np.pi * 5 * t) + 0.5 * np.sin(2 * np.pi * 20 * t)
            + 0.25 * np.sin(2 * np.pi * 60 * t) + 0.1 * rng.standard_normal(N))


def eeg_like(N, seed, fs=500.0, line_uv=0.0):
    rng = np.random.default_rng(seed)
    white = rng.standard_normal(N)
    spec = np.fft.rfft(white)
    f = np.fft.rfftfreq(N, 1.0 / fs)
    shape = np.zeros_like(f)
    shape[1:] = 1.0 / np.sqrt(f[1:])
    pink = np.fft.irfft(spec * shape, n=N)
    pink *= 30.0 / np.sqrt(np.mean(pink ** 2))
    t = np.arange(N) / fs
    dur = N / fs
    nb = rng.poisson(0.5 * dur)
    centers = rng.uniform(0.0, dur, nb)
    phases = rng.uniform(0.0, 2 * np.pi, nb)
    sigma = 0.5 / (2.0 * np.sqrt(2.0 * np.log(2.0)))   # FWHM 0.5 s
    alpha = np.zeros(N)
    for c, ph in zip(centers, phases):
        lo = max(0, int((c - 4 * sigma) * fs))
        hi = min(N, int((c + 4 * sigma) * fs) + 1)
        tt = t[lo:hi]
        alpha[lo:hi] += 20.0 * np.exp(-0.5 * ((tt - c) / sigma) ** 2) * np.sin(2 * np.pi * 10 * tt + ph)
    y = pink + alpha
    if line_uv:
        y = y + line_uv * np.sin(2 * np.pi * 50.0 * t + rng.uniform(0, 2 * np.pi))
    return y


def batch_lengths(C, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return [int(v) for v in rng.integers(lo, hi + 1, size=C)]
