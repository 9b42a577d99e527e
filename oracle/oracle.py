"""ctypes front end of the FP64 CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker.  The product path never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

STOP_FIXED = 0
STOP_RILLING = 1
BC_NATURAL = 0
BC_NOT_A_KNOT = 1

# windowed-solve geometry of the two paths: 64 lanes x B rows, H halo rows
# on each side (FP32 keeps 16: its truncation, ~1e-10 of the largest
# coefficient, is far below FP32 rounding)
GEOM_F64 = (8, 32)
GEOM_F32 = (8, 16)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from .build_oracle import build
            build()
        _lib = C.CDLL(LIB_PATH)
        _setup(_lib)
    return _lib


class _boundary:
    """Sets the oracle's spline boundary condition for the duration of a call."""

    def __init__(self, bc):
        self.bc = int(bc)

    def __enter__(self):
        lib().ocm_set_boundary(self.bc)

    def __exit__(self, *exc):
        lib().ocm_set_boundary(BC_NATURAL)


def _setup(L):
    dp = C.POINTER(C.c_double)
    L.ocm_set_boundary.argtypes = [C.c_int]
    L.ocm_set_boundary.restype = None
    ip = C.POINTER(C.c_int)
    L.ocm_extrema.argtypes = [dp, C.c_int, C.c_int, ip, dp]
    L.ocm_extrema.restype = C.c_int
    L.ocm_boundary_fix.argtypes = [C.c_int, ip, dp, C.c_int]
    L.ocm_spline_solve.argtypes = [C.c_int, ip, dp, C.c_int, C.c_int, dp]
    L.ocm_spline_solve_thomas.argtypes = [C.c_int, ip, dp, dp]
    L.ocm_solve_window.argtypes = [C.c_int, ip, dp, C.c_int, C.c_long, dp]
    L.ocm_envelope.argtypes = [C.c_int, C.c_int, ip, dp, dp, dp]
    L.ocm_sift_pass.argtypes = [dp, C.c_int, C.c_int, C.c_int, dp, ip, ip,
                                C.POINTER(C.c_int64), ip]
    L.ocm_rilling_continue.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int]
    L.ocm_rilling_continue.restype = C.c_int
    L.ocm_sum_canon.argtypes = [dp, C.c_long]
    L.ocm_sum_canon.restype = C.c_double
    L.ocm_var_canon.argtypes = [dp, C.c_long]
    L.ocm_var_canon.restype = C.c_double
    L.ocm_sift_mode.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, ip, C.c_int]
    L.ocm_sift_mode.restype = C.c_int
    L.ocm_iceemdan.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_double, dp, C.c_int, C.c_int, C.c_int, dp, ip, C.c_int, ip, dp]
    L.ocm_iceemdan.restype = C.c_int
    L.ocm_envelope_adjoint.argtypes = [C.c_int, C.c_int, ip, C.c_int, C.c_int, dp, dp]
    L.ocm_iceemdan_grad.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, dp,
                                    C.c_int, C.c_int, C.c_int, dp, dp, dp]
    L.ocm_iceemdan_grad.restype = C.c_int
    u64p = C.POINTER(C.c_uint64)
    L.ocm_philox4x64.argtypes = [u64p, u64p, u64p]
    L.ocm_noise.argtypes = [C.c_uint64, C.c_long, C.c_long, dp]


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def extrema(h, is_max):
    h = _f64(h)
    N = h.size
    pos = np.empty(N, np.int32)
    val = np.empty(N, np.float64)
    n = lib().ocm_extrema(_d(h), N, int(bool(is_max)), _i(pos), _d(val))
    return pos[:n].copy(), val[:n].copy()


def boundary_fix(pos, val, is_max):
    pos = np.ascontiguousarray(pos, np.int32)
    val = _f64(val).copy()
    lib().ocm_boundary_fix(pos.size, _i(pos), _d(val), int(bool(is_max)))
    return val


def spline_solve(pos, val, geom=GEOM_F64, bc=BC_NATURAL):
    pos = np.ascontiguousarray(pos, np.int32)
    val = _f64(val)
    c = np.empty(pos.size, np.float64)
    with _boundary(bc):
        lib().ocm_spline_solve(pos.size, _i(pos), _d(val), geom[0], geom[1], _d(c))
    return c


def spline_solve_thomas(pos, val, bc=BC_NATURAL):
    pos = np.ascontiguousarray(pos, np.int32)
    val = _f64(val)
    c = np.empty(pos.size, np.float64)
    with _boundary(bc):
        lib().ocm_spline_solve_thomas(pos.size, _i(pos), _d(val), _d(c))
    return c


def solve_window(pos, val, B, g0):
    pos = np.ascontiguousarray(pos, np.int32)
    val = _f64(val)
    x = np.empty(64 * B, np.float64)
    lib().ocm_solve_window(pos.size, _i(pos), _d(val), B, g0, _d(x))
    return x


def envelope(N, pos, val, c):
    pos = np.ascontiguousarray(pos, np.int32)
    env = np.empty(N, np.float64)
    lib().ocm_envelope(N, pos.size, _i(pos), _d(_f64(val)), _d(_f64(c)), _d(env))
    return env


def sift_pass(h, geom=GEOM_F64):
    h = _f64(h)
    N = h.size
    hn = np.empty(N, np.float64)
    nmax = C.c_int()
    nmin = C.c_int()
    cnt = C.c_int64()
    mark = C.c_int()
    lib().ocm_sift_pass(_d(h), N, geom[0], geom[1], _d(hn), C.byref(nmax), C.byref(nmin),
                        C.byref(cnt), C.byref(mark))
    return hn, nmax.value, nmin.value, cnt.value, mark.value


def rilling_continue(cnt_sx, any_mark, nmax, nmin, N):
    return bool(lib().ocm_rilling_continue(cnt_sx, any_mark, nmax, nmin, N))


def sum_canon(x):
    x = _f64(x)
    return lib().ocm_sum_canon(_d(x), x.size)


def var_canon(x):
    x = _f64(x)
    return lib().ocm_var_canon(_d(x), x.size)


def sift_mode(x, stop_mode, S_or_cap, geom=GEOM_F64, trace_passes=0):
    x = _f64(x)
    N = x.size
    mode = np.empty(N, np.float64)
    work = np.empty(3 * N, np.float64)
    trace = np.zeros(max(1, 2 * trace_passes), np.int32)
    subs = lib().ocm_sift_mode(_d(x), N, stop_mode, S_or_cap, geom[0], geom[1], _d(mode), _d(work),
                               _i(trace) if trace_passes else None, trace_passes)
    return mode, subs, trace[: 2 * trace_passes].reshape(-1, 2)


def iceemdan(y, noise, M, stop_mode=STOP_FIXED, S_or_cap=10, noise_cap=None, eps0=0.2,
             geom=GEOM_F64, nthreads=0, trace_passes=0, want_noise_modes=False, bc=BC_NATURAL):
    """Full ICEEMDAN on one channel.  noise: (I, N) standard normals.

    Returns dict(imfs=(M, N), trace=(2, M-1, P, I, 2) knot counts,
    stops=(2, M-1, I) subtraction counts, noise_modes=(M-1, I, N))."""
    y = _f64(y)
    noise = _f64(noise)
    I, N = noise.shape
    assert y.size == N
    if noise_cap is None:
        noise_cap = S_or_cap if stop_mode == STOP_FIXED else 2000
    imfs = np.empty((M, N), np.float64)
    P = trace_passes
    trace = np.zeros((2, M - 1, max(P, 1), I, 2), np.int32)
    stops = np.zeros((2, M - 1, I), np.int32)
    nm = np.empty((M - 1, I, N), np.float64) if want_noise_modes else None
    with _boundary(bc):
        rc = lib().ocm_iceemdan(_d(y), N, I, M, stop_mode, S_or_cap, noise_cap, eps0, _d(noise),
                                geom[0], geom[1], nthreads, _d(imfs), _i(trace) if P else None,
                                P, _i(stops), _d(nm) if nm is not None else None)
    if rc != 0:
        raise ValueError(f"oracle iceemdan failed: {rc}")
    out = dict(imfs=imfs, stops=stops)
    if P:
        out["trace"] = trace
    if nm is not None:
        out["noise_modes"] = nm
    return out


def philox4x64(ctr, key):
    c = (C.c_uint64 * 4)(*[int(v) & (2**64 - 1) for v in ctr])
    k = (C.c_uint64 * 2)(*[int(v) & (2**64 - 1) for v in key])
    o = (C.c_uint64 * 4)()
    lib().ocm_philox4x64(c, k, o)
    return [int(v) for v in o]


def noise(seed, I, N):
    out = np.empty((I, N), np.float64)
    lib().ocm_noise(C.c_uint64(seed), I, N, _d(out))
    return out


def envelope_adjoint(N, pos, g, fixed_lo=0, fixed_hi=0):
    pos = np.ascontiguousarray(pos, np.int32)
    out = np.zeros(N, np.float64)
    lib().ocm_envelope_adjoint(N, pos.size, _i(pos), int(fixed_lo), int(fixed_hi), _d(_f64(g)),
                               _d(out))
    return out


def iceemdan_grad(y, noise, M, grad_imfs, S=10, eps0=0.2, geom=GEOM_F64, nthreads=0):
    """dL/dy for L with dL/dIMF = grad_imfs (M x N), FIXED rule, frozen extrema.
    Returns (grad_y, imfs)."""
    y = _f64(y)
    noise = _f64(noise)
    I, N = noise.shape
    g = _f64(grad_imfs).reshape(M, N)
    gy = np.empty(N, np.float64)
    imfs = np.empty((M, N), np.float64)
    rc = lib().ocm_iceemdan_grad(_d(y), N, I, M, S, eps0, _d(noise), geom[0], geom[1], nthreads,
                                 _d(g), _d(gy), _d(imfs))
    if rc != 0:
        raise ValueError(f"oracle iceemdan_grad failed: {rc}")
    return gy, imfs
