"""Pins of the FP64 CPU oracle (CPU only).

Each check is chosen so that a plausible mistake in the oracle (dropped
term, wrong sign or index, transposed operand) fails it: independent
brute force, scipy's natural cubic spline, numpy statistics, numpy's
Philox, published known-answer vectors, and invariants of the
decomposition.
"""
import os

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.linalg import solve_banded

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def eeg(name):
    return np.fromfile(os.path.join(GOLD, name), dtype="<f4").astype(np.float64)


# ------------------------------------------------------------------ extrema

def brute_extrema(h, is_max):
    N = h.size
    out = []
    for p in range(N):
        if p == 0 or p == N - 1:
            out.append(p)
        elif is_max and h[p] > h[p - 1] and h[p] > h[p + 1]:
            out.append(p)
        elif not is_max and h[p] < h[p - 1] and h[p] < h[p + 1]:
            out.append(p)
    return np.array(out)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_extrema_brute_force(orc, seed):
    rng = np.random.default_rng(seed)
    for N in (4, 5, 17, 1000):
        h = rng.standard_normal(N)
        h[rng.integers(0, N, N // 5)] = 0.25            # plateaus / ties
        for is_max in (True, False):
            pos, val = orc.extrema(h, is_max)
            ref = brute_extrema(h, is_max)
            assert np.array_equal(pos, ref)
            assert np.array_equal(val, h[ref])


def test_extrema_nan_and_plateau(orc):
    h = np.array([0.0, 1.0, np.nan, 1.0, 2.0, 2.0, 1.0, 3.0, 0.0])
    pos, _ = orc.extrema(h, True)
    # NaN is never an extremum, a flat top (2,2) is not strict, endpoints forced
    assert list(pos) == [0, 7, 8]
    pos, _ = orc.extrema(h, False)
    assert list(pos) == [0, 6, 8]


# ------------------------------------------------------------- boundary fix

def test_boundary_fix_hand_cases(orc):
    pos = np.array([0, 2, 4, 7, 9, 12], np.int32)
    val = np.array([0.0, 1.0, 3.0, 2.0, 2.0, 1.0])
    # left: line through (2,1),(4,3) -> slope 1 -> at 0: -1 ; max keeps 0.0
    # right: line through (7,2),(9,2) -> 2 at 12 ; max replaces 1.0 by 2.0
    out = orc.boundary_fix(pos, val, True)
    assert out[0] == 0.0 and out[-1] == 2.0
    out = orc.boundary_fix(pos, val, False)
    assert out[0] == -1.0 and out[-1] == 1.0
    # n < 6: untouched
    out = orc.boundary_fix(pos[:5], val[:5], True)
    assert np.array_equal(out, val[:5])


# ------------------------------------------------------------------ spline

def natural_system(pos, val):
    n = pos.size
    ab = np.zeros((3, n))
    f = np.zeros(n)
    ab[1, :] = 1.0
    for g in range(1, n - 1):
        hl = pos[g] - pos[g - 1]
        hr = pos[g + 1] - pos[g]
        ab[2, g - 1] = hl          # sub-diagonal
        ab[1, g] = 2 * (hl + hr)
        ab[0, g + 1] = hr          # super-diagonal
        f[g] = 3 * ((val[g + 1] - val[g]) / hr - (val[g] - val[g - 1]) / hl)
    return ab, f


@pytest.mark.parametrize("src", ["white", "eeg", "sparse"])
def test_spline_solve_vs_scipy(orc, src):
    rng = np.random.default_rng(5)
    if src == "white":
        h = rng.standard_normal(30000)
    elif src == "eeg":
        h = eeg("eegSampleDataCH4.f32")
    else:
        h = np.cumsum(rng.standard_normal(20000))       # few, uneven knots
    pos, val = orc.extrema(h, True)
    c = orc.spline_solve(pos, val)
    ab, f = natural_system(pos, val)
    c_banded = solve_banded((1, 1), ab, f)
    cs = CubicSpline(pos.astype(float), val, bc_type="natural")
    scale = np.max(np.abs(c_banded))
    assert np.max(np.abs(c - c_banded)) <= 1e-13 * scale
    assert np.max(np.abs(c[:-1] - cs.c[1])) <= 1e-13 * scale
    assert c[0] == 0.0 and c[-1] == 0.0
    # residual of the tridiagonal system
    Ac = ab[1] * c
    Ac[1:] += ab[2, :-1] * c[:-1]
    Ac[:-1] += ab[0, 1:] * c[1:]
    assert np.max(np.abs(Ac - f)) <= 1e-12 * np.max(np.abs(f))


def test_spline_window_truncation_negligible(orc):
    """The windowed solve equals a global Thomas solve to round-off on EEG data
    (several windows: n >> 640)."""
    h = eeg("eegSampleDataCH1_244032.f32")
    for is_max in (True, False):
        pos, val = orc.extrema(h, is_max)
        assert pos.size > 10 * 640
        c = orc.spline_solve(pos, val)
        ct = orc.spline_solve_thomas(pos, val)
        assert np.max(np.abs(c - ct)) <= 1e-14 * np.max(np.abs(ct))


def test_spline_window_truncation_fp32_geometry(orc):
    """The FP32 path keeps 16 halo rows per side instead of 32.  Against a
    global Thomas solve the truncation is far below FP32 rounding (6e-8):
    measured 5.7e-11 (EEG maxima), 5.0e-11 (minima), 7.3e-11 (uniform
    spacing, the slowest decay of the coupling: 3.7x per row) and 5.6e-14
    (random spacings) of the largest coefficient; held to 1e-9.  With 32
    rows the same cases are at round-off (the test above)."""
    h = eeg("eegSampleDataCH1_244032.f32")
    cases = [orc.extrema(h, True), orc.extrema(h, False)]
    rng = np.random.default_rng(5)
    for hsp in (np.full(5000, 5), rng.integers(1, 300, 5000)):
        pos = np.concatenate([[0], np.cumsum(hsp)]).astype(np.int32)
        cases.append((pos, rng.standard_normal(pos.size)))
    for pos, val in cases:
        c = orc.spline_solve(pos, val, geom=orc.GEOM_F32)
        ct = orc.spline_solve_thomas(pos, val)
        err = np.max(np.abs(c - ct)) / np.max(np.abs(ct))
        assert 0 < err <= 1e-9, err


def test_spline_window_single_window_is_global(orc):
    rng = np.random.default_rng(9)
    pos = np.cumsum(rng.integers(1, 9, 300)).astype(np.int32)
    pos[0] = 0
    val = rng.standard_normal(pos.size)
    c = orc.spline_solve(pos, val)
    ct = orc.spline_solve_thomas(pos, val)
    assert np.max(np.abs(c - ct)) <= 1e-15 * np.max(np.abs(ct)) * 10


def test_envelope_vs_scipy_and_interpolation(orc):
    rng = np.random.default_rng(3)
    h = rng.standard_normal(5000)
    pos, val = orc.extrema(h, False)
    c = orc.spline_solve(pos, val)
    env = orc.envelope(h.size, pos, val, c)
    cs = CubicSpline(pos.astype(float), val, bc_type="natural")
    ref = cs(np.arange(h.size))
    assert np.max(np.abs(env - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert np.allclose(env[pos], val, rtol=0, atol=1e-12)
    assert env[0] == val[0]


def test_envelope_exact_on_linear_data(orc):
    pos = np.array([0, 3, 7, 8, 15, 20], np.int32)
    val = 2.0 + 0.5 * pos
    c = orc.spline_solve(pos, val)
    env = orc.envelope(21, pos, val, c)
    assert np.max(np.abs(c)) < 1e-15
    assert np.allclose(env, 2.0 + 0.5 * np.arange(21), rtol=0, atol=1e-13)


# --------------------------------------------------------------- sift pass

def test_sift_pass_independent(orc):
    """h_new = h - (U+L)/2 with U, L from scipy natural splines through the
    boundary-fixed extrema."""
    rng = np.random.default_rng(4)
    h = np.sin(np.arange(3000) / 9.0) + 0.3 * rng.standard_normal(3000)
    hn, nmax, nmin, cnt, mark = orc.sift_pass(h)
    env = []
    for is_max in (True, False):
        p = brute_extrema(h, is_max)
        v = h[p].copy()
        n = p.size
        if n >= 6:
            t = v[1] - (v[2] - v[1]) / (p[2] - p[1]) * (p[1] - p[0])
            tr = v[n - 2] + (v[n - 2] - v[n - 3]) / (p[n - 2] - p[n - 3]) * (p[n - 1] - p[n - 2])
            if is_max:
                v[0], v[-1] = max(v[0], t), max(v[-1], tr)
            else:
                v[0], v[-1] = min(v[0], t), min(v[-1], tr)
        env.append(CubicSpline(p.astype(float), v, bc_type="natural")(np.arange(h.size)))
    U, L = env
    assert (nmax, nmin) == (brute_extrema(h, True).size, brute_extrema(h, False).size)
    assert np.max(np.abs(hn - (h - (U + L) / 2))) < 1e-11
    sx = np.abs((U + L) / (U - L))
    assert abs(cnt - int(np.sum(sx > 0.05))) <= 2
    assert mark == int(np.any(sx > 0.5))


def test_sift_pass_few_knots_no_subtraction(orc):
    h = np.linspace(0.0, 1.0, 100) ** 2            # monotone: only the endpoints
    hn, nmax, nmin, cnt, mark = orc.sift_pass(h)
    assert nmax == 2 and nmin == 2
    assert np.array_equal(hn, h)


def test_rilling_truth_table(orc):
    N = 1000
    # frac = cnt/N > 0.05 continues regardless of the rest
    assert orc.rilling_continue(51, 0, 2, 2, N)
    assert not orc.rilling_continue(50, 0, 100, 100, N)
    # any(sx > 0.5) continues only with more than 6 extrema in total
    assert orc.rilling_continue(0, 1, 4, 3, N)
    assert not orc.rilling_continue(0, 1, 3, 3, N)
    assert not orc.rilling_continue(0, 0, 50, 50, N)


# ---------------------------------------------------------------- statistics

def test_canonical_sum_and_var(orc):
    rng = np.random.default_rng(8)
    for N in (2, 5, 1023, 1024, 1025, 244032):
        x = rng.standard_normal(N) * 3 + 1
        assert abs(orc.sum_canon(x) - np.sum(x)) <= 1e-12 * np.sum(np.abs(x))
        assert abs(orc.var_canon(x) - np.var(x, ddof=1)) <= 1e-12 * np.var(x)
    # exact order check against a hand restatement of the canonical order
    x = rng.standard_normal(3000)
    acc = np.zeros(1024)
    for p in range(3000):
        acc[p % 1024] += x[p]
    s = 512
    while s >= 1:
        acc[:s] = acc[:s] + acc[s:2 * s]
        s //= 2
    assert orc.sum_canon(x) == acc[0]


# ------------------------------------------------------------------- Philox

KAT = [  # Random123 philox4x64_10 known-answer vectors
    ([0, 0, 0, 0], [0, 0],
     [0x16554d9eca36314c, 0xdb20fe9d672d0fdc, 0xd7e772cee186176b, 0x7e68b68aec7ba23b]),
    ([2**64 - 1] * 4, [2**64 - 1] * 2,
     [0x87b092c3013fe90b, 0x438c3c67be8d0224, 0x9cc7d7c69cd777b6, 0xa09caebf594f0ba0]),
    ([0x243f6a8885a308d3, 0x13198a2e03707344, 0xa4093822299f31d0, 0x082efa98ec4e6c89],
     [0x452821e638d01377, 0xbe5466cf34e90c6c],
     [0xa528f45403e61d95, 0x38c72dbd566e9788, 0xa5a1610e72fd18b5, 0x57bd43b5e52b7fe6]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(orc, ctr, key, out):
    assert orc.philox4x64(ctr, key) == out


def test_philox_matches_numpy(orc):
    rng = np.random.default_rng(1)
    for _ in range(20):
        ctr = [int(v) for v in rng.integers(0, 2**63, 4, dtype=np.int64)]
        key = [int(v) for v in rng.integers(0, 2**63, 2, dtype=np.int64)]
        # numpy increments the counter before producing the first block
        c = list(ctr)
        c[0] = (c[0] - 1) % 2**64
        if ctr[0] == 0:
            c[1] = (c[1] - 1) % 2**64
        bg = np.random.Philox(key=np.array(key, dtype=np.uint64),
                              counter=np.array(c, dtype=np.uint64))
        assert [int(v) for v in bg.random_raw(4)] == orc.philox4x64(ctr, key)


def test_noise_statistics(orc):
    w = orc.noise(1, 8, 50000)
    assert abs(w.mean()) < 0.01
    assert abs(w.std() - 1.0) < 0.01
    # counter depends on (sample, realization) only: prefix-stable in N
    w2 = orc.noise(1, 8, 777)
    assert np.array_equal(w2, w[:, :777])
    assert not np.array_equal(orc.noise(2, 1, 100), w[:1, :100])


# ----------------------------------------------------------------- ICEEMDAN

def three_tone(N, seed=0):
    t = np.arange(N) / 1000.0
    rng = np.random.default_rng(seed)
    return (np.sin(2 * np.pi * 5 * t) + 0.5 * np.sin(2 * np.pi * 20 * t)
            + 0.25 * np.sin(2 * np.pi * 60 * t) + 0.1 * rng.standard_normal(N))


def test_iceemdan_completeness(orc):
    N, I, M = 4096, 20, 8
    x = three_tone(N)
    w = orc.noise(1, I, N)
    imfs = orc.iceemdan(x, w, M, S_or_cap=10)["imfs"]
    xt = x / np.sqrt(np.var(x, ddof=1))
    assert np.max(np.abs(imfs.sum(0) - xt)) < 1e-12
    # the 60 Hz tone lands in the first modes, the 5 Hz tone in a later one
    power = (imfs[:-1] ** 2).sum(1)
    assert np.argmax(power) >= 2


def test_iceemdan_eps0_zero_is_noise_free(orc):
    """eps0 = 0: every realization sifts x~ itself, so the result does not
    depend on I or on the noise (cudaICEEMDAN.cu:76,88 with eps0 = 0)."""
    N, M = 2048, 5
    x = three_tone(N, seed=3)
    a = orc.iceemdan(x, orc.noise(1, 3, N), M, S_or_cap=10, eps0=0.0)["imfs"]
    b = orc.iceemdan(x, orc.noise(7, 9, N), M, S_or_cap=10, eps0=0.0)["imfs"]
    assert np.max(np.abs(a - b)) < 1e-13


def test_iceemdan_rilling_stops(orc):
    N, I, M = 1000, 10, 3
    n = np.arange(N)
    x = np.sin(2 * np.pi * 0.065 * n) + np.where((n >= 500) & (n <= 748),
                                                  np.sin(2 * np.pi * 0.255 * (n - 500)), 0)
    r = orc.iceemdan(x, orc.noise(1234, I, N), M, stop_mode=orc.STOP_RILLING, S_or_cap=100)
    stops = r["stops"]
    assert stops.min() >= 0 and stops.max() <= 99
    imfs = r["imfs"]
    xt = x / np.sqrt(np.var(x, ddof=1))
    assert np.max(np.abs(imfs.sum(0) - xt)) < 1e-12


# ------------------------------------------------------ stage algebra pins
# Each check below fails on one plausible misreading of the stage algebra
# (cudaICEEMDAN.cu:41-90 kernels, :845-994 noise modes, :1003-1298 stages):
#   - a missing 1/std(E_1^i) at stage 0 (:76), or one std shared by all i
#   - a normalised E_{k+1} at stages k >= 1 (:88)
#   - the wrong noise-mode index at stage 0 or at stages k >= 1
#   - exporting h - m instead of h when the Rilling rule stops (:535-542)


def _alt_ramp(I, N, slopes, amps):
    """Realization i: slopes[i] * (p - N/2) + amps[i] * (-1)^(p+i).  With
    |amps| > |slopes| / 2 every interior sample is a strict extremum, and all
    values, differences and slopes are exact in binary floating point."""
    p = np.arange(N, dtype=np.float64)
    return np.stack([slopes[i] * (p - N // 2) + amps[i] * (-1.0) ** ((np.arange(N) + i) % 2)
                     for i in range(I)])


def test_noise_modes_of_exact_alternation_and_ramp(orc):
    """Noise = ramp + alternation: the first sift pass meets two straight
    envelopes (the endpoint fix puts the end knots on the lines), so
    E_1 = the alternation, E_2 = the ramp (no interior extrema: nothing is
    subtracted) and E_k = 0 for k >= 3, all exactly."""
    N, I, M = 512, 3, 5
    slopes, amps = [1.0, -0.5, 2.0], [1.0, 0.75, 1.5]
    w = _alt_ramp(I, N, slopes, amps)
    r = orc.iceemdan(three_tone(N), w, M, S_or_cap=10, want_noise_modes=True)
    E = r["noise_modes"]
    ramp = _alt_ramp(I, N, slopes, [0.0] * I)
    assert np.array_equal(E[0], w - ramp)
    assert np.array_equal(E[1], ramp)
    assert np.all(E[2:] == 0.0)


def test_stage_noise_normalisation_and_index(orc):
    """Stage 0 adds eps0 E_1^i / std(E_1^i), stage k >= 1 adds the raw
    eps0 std(r_k) E_{k+1}^i:
      (a) scaling one realization's E_1 by 2 (its alternation) changes
          nothing: stage 0 normalises per realization, later stages do not
          see E_1;
      (b) changing E_2 only (the ramps' slopes) keeps row 0 and changes
          row 1: stage 0 uses E_1, stage 1 uses E_2;
      (c) scaling the whole noise by 2 (exact in floating point) keeps row 0
          bit for bit and changes rows >= 1: E_{k+1} is not normalised at
          k >= 1."""
    N, I, M = 512, 3, 5
    x = three_tone(N, seed=4)
    base = orc.iceemdan(x, _alt_ramp(I, N, [1.0, -0.5, 2.0], [1.0, 0.75, 1.5]), M,
                        S_or_cap=10)["imfs"]
    amp2 = orc.iceemdan(x, _alt_ramp(I, N, [1.0, -0.5, 2.0], [2.0, 0.75, 1.5]), M,
                        S_or_cap=10)["imfs"]
    assert np.array_equal(amp2, base)
    slope = orc.iceemdan(x, _alt_ramp(I, N, [0.5, -0.25, 1.0], [1.0, 0.75, 1.5]), M,
                         S_or_cap=10)["imfs"]
    assert np.array_equal(slope[0], base[0])
    assert not np.array_equal(slope[1], base[1])
    w = orc.noise(3, 6, 1500)
    x = three_tone(1500, seed=5)
    one = orc.iceemdan(x, w, M, S_or_cap=10)["imfs"]
    two = orc.iceemdan(x, 2.0 * w, M, S_or_cap=10)["imfs"]
    assert np.array_equal(one[0], two[0])
    for k in range(1, M):
        assert not np.array_equal(one[k], two[k]), k
    # and the noise does enter stage 0: eps0 = 0 gives another row 0
    zero = orc.iceemdan(x, w, M, S_or_cap=10, eps0=0.0)["imfs"]
    assert np.max(np.abs(zero[0] - one[0])) > 1e-3


@pytest.mark.parametrize("noise_cap", [1, 2000])
def test_rilling_cap_one_exports_unsifted(orc, noise_cap):
    """RILLING with a cap of one pass exports each realization's input
    unsubtracted (cudaICEEMDAN.cu:535-542): r_1 = mean(x_i - x_i) = 0, so
    the rows are exactly [x~, 0, ..., 0]."""
    N, I, M = 800, 5, 4
    x = three_tone(N, seed=6)
    r = orc.iceemdan(x, orc.noise(4, I, N), M, stop_mode=orc.STOP_RILLING, S_or_cap=1,
                     noise_cap=noise_cap)
    imfs = r["imfs"]
    xt = x / np.sqrt(np.var(x, ddof=1))
    assert np.max(np.abs(imfs[0] - xt)) <= 1e-14 * np.max(np.abs(xt))
    assert np.all(imfs[1:] == 0.0)
    assert np.all(r["stops"][1] == 0)


def test_box_muller_pairs(orc):
    """The Gaussian stream is Box-Muller on numpy's Philox4x64-10 words:
    z0 = sqrt(-2 ln u0) cos(2 pi u1), z1 = sqrt(-2 ln u0) sin(2 pi u1) (and
    the same on u2, u3), u = ((bits >> 11) + 1/2) 2^-53; the pairs are
    uncorrelated and each component passes a Kolmogorov-Smirnov test."""
    from scipy import stats
    I, N, seed = 3, 40000, 9
    w = orc.noise(seed, I, N)
    for i in range(I):
        # numpy increments its counter before each block: start one below
        # (0, i, 0, 0), borrowing through the words
        start = [2**64 - 1, i - 1, 0, 0] if i else [2**64 - 1] * 4
        bg = np.random.Philox(key=np.array([seed, 0], dtype=np.uint64),
                              counter=np.array(start, dtype=np.uint64))
        bits = bg.random_raw(N).reshape(-1, 4)
        u = ((bits >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
        r0 = np.sqrt(-2.0 * np.log(u[:, 0]))
        r1 = np.sqrt(-2.0 * np.log(u[:, 2]))
        ref = np.stack([r0 * np.cos(2 * np.pi * u[:, 1]), r0 * np.sin(2 * np.pi * u[:, 1]),
                        r1 * np.cos(2 * np.pi * u[:, 3]), r1 * np.sin(2 * np.pi * u[:, 3])],
                       axis=1).reshape(-1)
        assert np.max(np.abs(w[i] - ref)) < 1e-13
    z = w.reshape(-1, 4)
    n = z.shape[0]
    for a, b in ((0, 1), (2, 3), (0, 2)):
        assert abs(np.corrcoef(z[:, a], z[:, b])[0, 1]) < 4.0 / np.sqrt(n)
    for j in range(4):
        assert stats.kstest(z[:, j], "norm").pvalue > 1e-3, j
    assert not np.array_equal(z[:, 0], z[:, 1])


# ---------------------------------------------------------- not-a-knot

@pytest.mark.parametrize("n", [4, 5, 9, 40, 3000])
def test_not_a_knot_matches_scipy(orc, n):
    """c_k (half the second derivative) of the not-a-knot spline equals
    scipy's CubicSpline(bc_type='not-a-knot'); the windowed solve (several
    windows at n = 3000) equals the global one."""
    rng = np.random.default_rng(n)
    pos = np.concatenate([[0], np.cumsum(rng.integers(1, 40, n - 1))]).astype(np.int32)
    val = rng.standard_normal(n)
    cs = CubicSpline(pos.astype(np.float64), val, bc_type="not-a-knot")
    ref = cs(pos.astype(np.float64), 2) / 2.0
    ct = orc.spline_solve_thomas(pos, val, bc=orc.BC_NOT_A_KNOT)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(ct - ref)) <= 1e-11 * scale
    cw = orc.spline_solve(pos, val, bc=orc.BC_NOT_A_KNOT)
    assert np.max(np.abs(cw - ct)) <= 1e-12 * scale
    # the third derivative is continuous at x_1 and x_{n-2}
    d = np.diff(ct) / (3.0 * np.diff(pos))
    assert abs(d[0] - d[1]) <= 1e-9 * np.max(np.abs(d))
    assert abs(d[-1] - d[-2]) <= 1e-9 * np.max(np.abs(d))
    # the setting does not leak into later natural solves
    cn = orc.spline_solve_thomas(pos, val)
    assert cn[0] == 0.0 and cn[-1] == 0.0


def test_iceemdan_not_a_knot_completeness(orc):
    N, I, M = 2048, 6, 5
    x = three_tone(N, seed=5)
    w = orc.noise(2, I, N)
    nak = orc.iceemdan(x, w, M, S_or_cap=10, bc=orc.BC_NOT_A_KNOT)["imfs"]
    nat = orc.iceemdan(x, w, M, S_or_cap=10)["imfs"]
    xt = x / np.sqrt(np.var(x, ddof=1))
    assert np.max(np.abs(nak.sum(0) - xt)) < 1e-12
    assert not np.array_equal(nak, nat)


# ---------------------------------------------------------------- adjoint

def _frozen_envelope(orc, N, pos, flo, fhi):
    """The envelope as a linear map of the samples at the knot positions,
    with the endpoint-fix branch held fixed; built from the oracle's
    forward pieces (scipy pins those), independent of the adjoint code."""
    def env(h):
        v = h[pos].copy()
        if flo:
            r = (pos[1] - pos[0]) / (pos[2] - pos[1])
            v[0] = v[1] * (1 + r) - v[2] * r
        if fhi:
            r = (pos[-1] - pos[-2]) / (pos[-2] - pos[-3])
            v[-1] = v[-2] * (1 + r) - v[-3] * r
        return orc.envelope(N, pos, v, orc.spline_solve_thomas(pos, v))
    return env


@pytest.mark.parametrize("seed,is_max", [(0, True), (1, False), (5, True), (8, False)])
def test_envelope_adjoint_is_jacobian_transpose(orc, seed, is_max):
    """ocm_envelope_adjoint(g) == J^T g with J the dense Jacobian of the
    frozen-knot envelope (assembled column by column from the forward)."""
    rng = np.random.default_rng(seed)
    N = 240
    h = rng.standard_normal(N)
    pos, val = orc.extrema(h, is_max)
    fixed = orc.boundary_fix(pos, val, is_max)
    flo, fhi = int(fixed[0] != val[0]), int(fixed[-1] != val[-1])
    env = _frozen_envelope(orc, N, pos, flo, fhi)
    J = np.stack([env(np.eye(N)[q]) for q in range(N)], axis=1)
    g = rng.standard_normal(N)
    adj = orc.envelope_adjoint(N, pos, g, flo, fhi)
    ref = J.T @ g
    assert np.max(np.abs(adj - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_envelope_adjoint_both_fix_branches(orc):
    """A signal whose ends both get extrapolated (rising ramps into each
    end): both endpoint-fix adjoint branches are exercised."""
    N = 200
    n = np.arange(N)
    h = np.sin(2 * np.pi * n / 23.0) + 0.02 * (n - N / 2) ** 2 / N
    found = False
    for is_max in (True, False):
        pos, val = orc.extrema(h, is_max)
        fixed = orc.boundary_fix(pos, val, is_max)
        flo, fhi = int(fixed[0] != val[0]), int(fixed[-1] != val[-1])
        if not (flo and fhi):
            continue
        found = True
        env = _frozen_envelope(orc, N, pos, flo, fhi)
        J = np.stack([env(np.eye(N)[q]) for q in range(N)], axis=1)
        g = np.cos(n / 7.0)
        adj = orc.envelope_adjoint(N, pos, g, flo, fhi)
        assert np.max(np.abs(adj - J.T @ g)) <= 1e-12 * np.max(np.abs(J.T @ g))
    assert found


@pytest.mark.parametrize("eps0", [0.2, 0.0])
def test_iceemdan_grad_matches_finite_differences(orc, eps0):
    """dL/dy of L = sum_k w_k |IMF_k|^2 against central differences
    (h = 1e-5 keeps every extrema set unchanged; the error of the
    difference quotient is then O(h^2))."""
    rng = np.random.default_rng(11)
    N, I, M, S = 1024, 6, 5, 10
    t = np.arange(N) / 1000.0
    x = np.sin(2 * np.pi * 5 * t) + 0.5 * np.sin(2 * np.pi * 40 * t) + 0.1 * rng.standard_normal(N)
    w = orc.noise(1, I, N)
    wk = rng.uniform(0, 1, M)

    def loss(xx):
        im = orc.iceemdan(xx, w, M, S_or_cap=S, eps0=eps0)["imfs"]
        return float(np.sum(wk * np.sum(im ** 2, 1))), im

    _, im0 = loss(x)
    gy, im = orc.iceemdan_grad(x, w, M, 2 * wk[:, None] * im0, S=S, eps0=eps0)
    assert np.array_equal(im, im0)
    for _ in range(3):
        v = rng.standard_normal(N)
        v /= np.linalg.norm(v)
        hh = 1e-5
        fd = (loss(x + hh * v)[0] - loss(x - hh * v)[0]) / (2 * hh)
        assert abs(fd - gy @ v) <= 1e-6 * abs(fd), (fd, gy @ v)


def test_golden_fixtures(orc):
    """The oracle reproduces its committed outputs bit for bit."""
    path = os.path.join(GOLD, "oracle_fixtures.npz")
    if not os.path.exists(path):
        pytest.skip("fixtures not generated")
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_fixtures",
                                                  os.path.join(GOLD, "make_fixtures.py"))
    mf = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mf)
    CASES, run_case = mf.CASES, mf.run_case
    data = np.load(path, allow_pickle=False)
    for name in CASES:
        got = run_case(orc, name)
        for k, v in got.items():
            assert np.array_equal(data[f"{name}/{k}"], v), (name, k)
