"""MI355X-native Improved CEEMDAN (host side of the HIP library).

Python mirror of the reference interface (cudaICEEMDAN.cuh:25-26,
statistics.cuh:25-44) over the C ABI of libiceemdan_mi355x.so
(include/iceemdan_mi355x.h).  PyTorch provides device memory and streams;
every computation runs in the library's HIP kernels.  There is no CPU
fallback: if the library is missing or no GPU is present the calls raise.

The package directory name contains a hyphen; import it with
``importlib.import_module("ceemdan-gpu_amd")``.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libiceemdan_mi355x.so")

ICM_OK = 0
DTYPE_F32 = 0
DTYPE_F64 = 1
STOP_FIXED = 0
STOP_RILLING = 1
BOUNDARY_NATURAL = 0
BOUNDARY_NOT_A_KNOT = 1
PHASE_NOISE = 0
PHASE_SIGNAL = 1

_lib = None


class ICMError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = lib().icm_status_string(status).decode()
        if status in (-2, -3):
            detail = lib().icm_last_error().decode()
            if detail:
                msg += f" [{detail}]"
        super().__init__(f"{what}: {msg} ({status})" if what else f"{msg} ({status})")


class icm_config(C.Structure):
    _fields_ = [("dtype", C.c_int), ("num_channels", C.c_int),
                ("lengths", C.POINTER(C.c_int64)), ("num_noise", C.c_int),
                ("num_imfs", C.c_int), ("stop_mode", C.c_int), ("max_sifts", C.c_int),
                ("noise_max_sifts", C.c_int), ("noise_strength", C.c_double),
                ("seed", C.c_uint64), ("use_graph", C.c_int), ("boundary", C.c_int),
                ("keep_records", C.c_int)]


class icm_plan_info(C.Structure):
    _fields_ = [("device_bytes", C.c_int64), ("noise_series", C.c_int64),
                ("signal_series", C.c_int64), ("noise_samples", C.c_int64),
                ("signal_samples", C.c_int64), ("tiles_noise", C.c_int64),
                ("tiles_signal", C.c_int64), ("tile", C.c_int32),
                ("solve_rows_per_lane", C.c_int32), ("solve_halo", C.c_int32)]


class icm_timing(C.Structure):
    _fields_ = [("noise_ms", C.c_double), ("signal_ms", C.c_double), ("total_ms", C.c_double),
                ("backward_ms", C.c_double)]


# exported C symbols and their signatures (kept in step with include/*.h)
_SIGS = {
    "icm_status_string": (C.c_char_p, [C.c_int]),
    "icm_last_error": (C.c_char_p, []),
    "icm_config_default": (C.c_int, [C.POINTER(icm_config)]),
    "icm_validate_config": (C.c_int, [C.POINTER(icm_config)]),
    "icm_plan_bytes": (C.c_int, [C.POINTER(icm_config), C.POINTER(C.c_int64)]),
    "icm_plan_create": (C.c_int, [C.POINTER(icm_config), C.POINTER(C.c_void_p)]),
    "icm_plan_destroy": (C.c_int, [C.c_void_p]),
    "icm_plan_get_info": (C.c_int, [C.c_void_p, C.POINTER(icm_plan_info)]),
    "icm_forward": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                              C.c_void_p, C.c_void_p]),
    "icm_last_timing": (C.c_int, [C.c_void_p, C.POINTER(icm_timing)]),
    "icm_set_autotune": (C.c_int, [C.c_void_p, C.c_int]),
    "icm_get_tuning": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.c_int64]),
    "icm_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.c_int64]),
    "icm_tuning_candidate": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "icm_get_tuning_left": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "icm_get_graph_status": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "icm_set_overlap": (C.c_int, [C.c_void_p, C.c_int]),
    "icm_set_trace": (C.c_int, [C.c_void_p, C.c_int]),
    "icm_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "icm_get_kernel_stats": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "icm_get_trace": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.c_int64]),
    "icm_get_trace_hash": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint64), C.c_int64]),
    "icm_get_stops": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.c_int64]),
    "icm_get_noise_modes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "icm_noise": (C.c_int, [C.c_int, C.c_uint64, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p]),
    "icm_sift_pass": (C.c_int, [C.c_int, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p,
                                C.c_void_p, C.c_void_p]),
    "icm_sift_pass_static": (C.c_int, [C.c_int, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p,
                                       C.c_void_p, C.c_void_p]),
    "icm_backward": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                               C.c_void_p]),
    "icm_forward_begin": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p]),
    "icm_forward_stages": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_void_p]),
    "icm_forward_finish": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_void_p]),
    "icm_state_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "icm_state_save": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "icm_state_load": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "icm_stats_mean": (C.c_int, [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "icm_stats_variance": (C.c_int, [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                     C.c_void_p, C.c_void_p]),
    "icm_stats_multiply": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                                     C.c_void_p]),
    "icm_iceemdanS": (C.c_double, [C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_float]),
    "icm_iceemdanD": (C.c_double, [C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_double]),
}


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status, what=""):
    if status != ICM_OK:
        raise ICMError(status, what)


def _torch():
    import torch
    return torch


def _dtype_code(dtype):
    torch = _torch()
    if dtype == torch.float64:
        return DTYPE_F64
    if dtype == torch.float32:
        return DTYPE_F32
    raise TypeError(f"unsupported dtype {dtype}")


def _stream_ptr(device):
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _require_cuda(*ts):
    for t in ts:
        if not t.is_cuda:
            raise ValueError("tensors must live on the GPU")
        if not t.is_contiguous():
            raise ValueError("tensors must be contiguous")


def _check_rows(bufs, lengths, rows, dtype, what):
    """Caller-supplied per-channel buffers: one per channel, the plan's dtype,
    rows * N_c elements each, on the GPU and contiguous (the library writes
    exactly that many elements through each pointer)."""
    if len(bufs) != len(lengths):
        raise ValueError(f"{what}: {len(bufs)} buffers for {len(lengths)} channels")
    for b, n in zip(bufs, lengths):
        _require_cuda(b)
        if b.dtype != dtype:
            raise ValueError(f"{what}: dtype {b.dtype} differs from the plan's {dtype}")
        if b.numel() != rows * n:
            raise ValueError(f"{what}: {b.numel()} elements where {rows} x {n} are needed")


def _check_noise(noise, num_noise, lengths, dtype):
    """Caller noise: [I, N] of the plan's dtype; all channels must share N."""
    _require_cuda(noise)
    if noise.dtype != dtype:
        raise ValueError("noise dtype differs from the plan")
    if len(set(lengths)) != 1:
        raise ValueError("caller noise needs channels of one length")
    if noise.numel() != num_noise * lengths[0]:
        raise ValueError(f"noise has {noise.numel()} elements where "
                         f"{num_noise} x {lengths[0]} are needed")


def make_config(lengths, num_noise, num_imfs, dtype_code=DTYPE_F64, stop_mode=STOP_FIXED,
                max_sifts=10, noise_max_sifts=None, noise_strength=0.2, seed=1234, use_graph=1,
                boundary=BOUNDARY_NATURAL, keep_records=False):
    cfg = icm_config()
    lib().icm_config_default(C.byref(cfg))
    lens = (C.c_int64 * len(lengths))(*[int(v) for v in lengths])
    cfg.dtype = dtype_code
    cfg.num_channels = len(lengths)
    cfg.lengths = C.cast(lens, C.POINTER(C.c_int64))
    cfg.num_noise = int(num_noise)
    cfg.num_imfs = int(num_imfs)
    cfg.stop_mode = int(stop_mode)
    cfg.max_sifts = int(max_sifts)
    if noise_max_sifts is None:
        noise_max_sifts = max_sifts if stop_mode == STOP_FIXED else 2000
    cfg.noise_max_sifts = int(noise_max_sifts)
    cfg.noise_strength = float(noise_strength)
    cfg.seed = int(seed)
    cfg.use_graph = int(use_graph)
    cfg.boundary = int(boundary)
    cfg.keep_records = 1 if keep_records else 0
    return cfg, lens


def validate_config(**kw):
    """Host-only argument validation (no device work); returns the status."""
    cfg, _keep = make_config(**kw)
    return lib().icm_validate_config(C.byref(cfg))


def plan_bytes(**kw):
    cfg, _keep = make_config(**kw)
    b = C.c_int64()
    _check(lib().icm_plan_bytes(C.byref(cfg), C.byref(b)), "icm_plan_bytes")
    return b.value


class Plan:
    """A reusable ICEEMDAN plan: all channels, realizations and scratch.

    Mirrors one call of iceemdanS/iceemdanD (cudaICEEMDAN.cu:619/:1340),
    generalised to many channels of uneven length."""

    def __init__(self, lengths, num_noise, num_imfs, dtype=None, stop_mode=STOP_FIXED,
                 max_sifts=10, noise_max_sifts=None, noise_strength=0.2, seed=1234,
                 use_graph=True, device=0, boundary=BOUNDARY_NATURAL, keep_records=False):
        torch = _torch()
        if dtype is None:
            dtype = torch.float64
        self.dtype = dtype
        self.lengths = [int(v) for v in lengths]
        self.num_noise = int(num_noise)
        self.num_imfs = int(num_imfs)
        self.stop_mode = stop_mode
        self.device = torch.device("cuda", device) if isinstance(device, int) else device
        cfg, self._lens = make_config(self.lengths, num_noise, num_imfs, _dtype_code(dtype),
                                      stop_mode, max_sifts, noise_max_sifts, noise_strength,
                                      seed, 1 if use_graph else 0, boundary, keep_records)
        self._cfg = cfg
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _check(lib().icm_plan_create(C.byref(cfg), C.byref(h)), "icm_plan_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().icm_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        inf = icm_plan_info()
        _check(lib().icm_plan_get_info(self._h, C.byref(inf)))
        return {f: getattr(inf, f) for f, _ in icm_plan_info._fields_}

    def set_trace(self, passes):
        _check(lib().icm_set_trace(self._h, int(passes)), "icm_set_trace")

    def trace(self, phase):
        import numpy as np
        inf = self.info()
        nser = inf["noise_series"] if phase == PHASE_NOISE else inf["signal_series"]
        P = self._trace_passes
        out = np.zeros((self.num_imfs - 1, P, nser, 2), np.int32)
        _check(lib().icm_get_trace(self._h, phase, out.ctypes.data_as(C.POINTER(C.c_int32)),
                                   out.size), "icm_get_trace")
        return out

    def trace_hash(self, phase):
        """[M-1, passes, series] uint64 digests of the knot sets of every
        traced pass input (equal digests: equal extrema index sets)."""
        import numpy as np
        inf = self.info()
        nser = inf["noise_series"] if phase == PHASE_NOISE else inf["signal_series"]
        out = np.zeros((self.num_imfs - 1, self._trace_passes, nser), np.uint64)
        _check(lib().icm_get_trace_hash(self._h, phase,
                                        out.ctypes.data_as(C.POINTER(C.c_uint64)), out.size),
               "icm_get_trace_hash")
        return out

    def stops(self, phase):
        import numpy as np
        inf = self.info()
        nser = inf["noise_series"] if phase == PHASE_NOISE else inf["signal_series"]
        out = np.zeros((self.num_imfs - 1, nser), np.int32)
        _check(lib().icm_get_stops(self._h, phase, out.ctypes.data_as(C.POINTER(C.c_int32)),
                                   out.size), "icm_get_stops")
        return out

    def enable_trace(self, passes):
        self._trace_passes = int(passes)
        self.set_trace(passes)

    def forward(self, signals, out=None, noise=None):
        """signals: list of 1-D device tensors (one per channel).
        Returns a list of [num_imfs, N_c] tensors."""
        torch = _torch()
        if isinstance(signals, torch.Tensor):
            signals = [signals] if signals.dim() == 1 else list(signals)
        if len(signals) != len(self.lengths):
            raise ValueError("number of channels differs from the plan")
        for s, n in zip(signals, self.lengths):
            if s.numel() != n or s.dtype != self.dtype:
                raise ValueError("channel length/dtype differs from the plan")
            _require_cuda(s)
        if out is None:
            out = [torch.empty((self.num_imfs, n), dtype=self.dtype, device=signals[0].device)
                   for n in self.lengths]
        _check_rows(out, self.lengths, self.num_imfs, self.dtype, "forward out")
        if noise is not None:
            _check_noise(noise, self.num_noise, self.lengths, self.dtype)
        sp = (C.c_void_p * len(signals))(*[s.data_ptr() for s in signals])
        op = (C.c_void_p * len(out))(*[o.data_ptr() for o in out])
        st = _stream_ptr(signals[0].device)
        _check(lib().icm_forward(self._h, sp, op,
                                 C.c_void_p(noise.data_ptr()) if noise is not None else None, st),
               "icm_forward")
        self.forward_count = getattr(self, "forward_count", 0) + 1
        return out

    # ---- staged forward with stage-boundary checkpoints
    def _signals(self, signals):
        torch = _torch()
        if isinstance(signals, torch.Tensor):
            signals = [signals] if signals.dim() == 1 else list(signals)
        if len(signals) != len(self.lengths):
            raise ValueError("number of channels differs from the plan")
        for s, n in zip(signals, self.lengths):
            if s.numel() != n or s.dtype != self.dtype:
                raise ValueError("channel length/dtype differs from the plan")
            _require_cuda(s)
        return signals

    def forward_begin(self, signals, noise=None):
        """Inputs, noise modes and stage-0 statistics; returns stage 0."""
        signals = self._signals(signals)
        if noise is not None:
            _check_noise(noise, self.num_noise, self.lengths, self.dtype)
        sp = (C.c_void_p * len(signals))(*[s.data_ptr() for s in signals])
        _check(lib().icm_forward_begin(self._h, sp,
                                       C.c_void_p(noise.data_ptr()) if noise is not None else None,
                                       _stream_ptr(self.device)), "icm_forward_begin")
        return 0

    def forward_stages(self, count):
        """Runs up to `count` CEEMDAN stages; returns the next stage."""
        st = C.c_int()
        _check(lib().icm_forward_stages(self._h, int(count), C.byref(st), _stream_ptr(self.device)),
               "icm_forward_stages")
        return st.value

    def forward_finish(self, out=None):
        torch = _torch()
        if out is None:
            out = [torch.empty((self.num_imfs, n), dtype=self.dtype, device=self.device)
                   for n in self.lengths]
        _check_rows(out, self.lengths, self.num_imfs, self.dtype, "forward_finish out")
        op = (C.c_void_p * len(out))(*[o.data_ptr() for o in out])
        _check(lib().icm_forward_finish(self._h, op, _stream_ptr(self.device)), "icm_forward_finish")
        self.forward_count = getattr(self, "forward_count", 0) + 1
        return out

    def save_state(self):
        """Host copy (numpy uint8) of the state between two stages."""
        import numpy as np
        n = C.c_int64()
        _check(lib().icm_state_bytes(self._h, C.byref(n)), "icm_state_bytes")
        buf = np.empty(n.value, np.uint8)
        _check(lib().icm_state_save(self._h, C.c_void_p(buf.ctypes.data), n.value,
                                    _stream_ptr(self.device)), "icm_state_save")
        return buf

    def load_state(self, buf, noise=None):
        """Resumes a saved state; returns its stage."""
        import numpy as np
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        if noise is not None:
            _check_noise(noise, self.num_noise, self.lengths, self.dtype)
        _check(lib().icm_state_load(self._h, C.c_void_p(buf.ctypes.data), buf.size,
                                    C.c_void_p(noise.data_ptr()) if noise is not None else None,
                                    _stream_ptr(self.device)), "icm_state_load")
        return self.forward_stages(0)

    def backward(self, grad_imfs, out=None):
        """dL/dx of the last forward from dL/dIMF (list of [num_imfs, N_c]
        device tensors), FIXED stop rule, extrema sets frozen at the
        forward's.  Returns a list of [N_c] tensors."""
        torch = _torch()
        if isinstance(grad_imfs, torch.Tensor):
            grad_imfs = [grad_imfs] if grad_imfs.dim() == 2 else list(grad_imfs)
        if len(grad_imfs) != len(self.lengths):
            raise ValueError("number of channels differs from the plan")
        gs = []
        for g, n in zip(grad_imfs, self.lengths):
            g = g.contiguous()
            if g.numel() != n * self.num_imfs or g.dtype != self.dtype:
                raise ValueError("gradient shape/dtype differs from the plan")
            _require_cuda(g)
            gs.append(g)
        if out is None:
            out = [torch.empty(n, dtype=self.dtype, device=gs[0].device) for n in self.lengths]
        _check_rows(out, self.lengths, 1, self.dtype, "backward out")
        gp = (C.c_void_p * len(gs))(*[g.data_ptr() for g in gs])
        op = (C.c_void_p * len(out))(*[o.data_ptr() for o in out])
        _check(lib().icm_backward(self._h, gp, op, _stream_ptr(gs[0].device)), "icm_backward")
        return out

    # ---- autotuner of the sift pass
    def set_autotune(self, forwards):
        """Time the sift candidates over the next `forwards` forwards (0: off,
        default geometry)."""
        _check(lib().icm_set_autotune(self._h, int(forwards)), "icm_set_autotune")

    def set_overlap(self, on):
        """FIXED mode: the noise-mode phase beside the CEEMDAN phase on a
        second stream (default), or one after the other."""
        _check(lib().icm_set_overlap(self._h, 1 if on else 0), "icm_set_overlap")

    def graph_status(self):
        """1: the phases replay captured hipGraphs; 0: eager; -1: a capture
        failed (the plan runs eagerly)."""
        n = C.c_int()
        _check(lib().icm_get_graph_status(self._h, C.byref(n)), "icm_get_graph_status")
        return n.value

    def tuning_left(self):
        """Forwards that will still time the sift candidates."""
        n = C.c_int()
        _check(lib().icm_get_tuning_left(self._h, C.byref(n)), "icm_get_tuning_left")
        return n.value

    def tuning(self, phase):
        """Chosen candidate per stage of a phase (-1 = default geometry)."""
        import numpy as np
        out = np.zeros(self.num_imfs - 1, np.int32)
        _check(lib().icm_get_tuning(self._h, phase, out.ctypes.data_as(C.POINTER(C.c_int32)),
                                    out.size), "icm_get_tuning")
        return out

    def set_tuning(self, phase, cands):
        import numpy as np
        a = np.ascontiguousarray(cands, dtype=np.int32)
        _check(lib().icm_set_tuning(self._h, phase, a.ctypes.data_as(C.POINTER(C.c_int32)),
                                    a.size), "icm_set_tuning")

    def set_profiling(self, on):
        _check(lib().icm_set_profiling(self._h, 1 if on else 0), "icm_set_profiling")

    def kernel_stats(self):
        """{kind: (total_ms, launches, algorithmic_bytes)} of the last forward."""
        out = {}
        for k, name in ((0, "init"), (1, "solve"), (2, "sift")):
            ms, n, b = C.c_double(), C.c_int64(), C.c_int64()
            _check(lib().icm_get_kernel_stats(self._h, k, C.byref(ms), C.byref(n), C.byref(b)))
            out[name] = (ms.value, n.value, b.value)
        return out

    def timing(self):
        t = icm_timing()
        _check(lib().icm_last_timing(self._h, C.byref(t)))
        return {"noise_ms": t.noise_ms, "signal_ms": t.signal_ms, "total_ms": t.total_ms,
                "backward_ms": t.backward_ms}

    def noise_modes(self):
        torch = _torch()
        inf = self.info()
        n = (self.num_imfs - 1) * inf["noise_samples"]
        out = torch.empty(n, dtype=self.dtype, device=self.device)
        _check(lib().icm_get_noise_modes(self._h, C.c_void_p(out.data_ptr()), n,
                                         _stream_ptr(self.device)), "icm_get_noise_modes")
        return out


def tuning_candidate(cand):
    """(tiles per segment or 0 = default, small tables 1 / full 0 / default -1)."""
    seg, small = C.c_int(), C.c_int()
    _check(lib().icm_tuning_candidate(int(cand), C.byref(seg), C.byref(small)),
           "icm_tuning_candidate")
    return seg.value, small.value


NUM_TUNING_CANDIDATES = 8


def noise(num_noise, length, seed=1234, dtype=None, device=0):
    """[I, N] standard normals from the library's Philox4x64-10 generator."""
    torch = _torch()
    dtype = dtype or torch.float64
    out = torch.empty((num_noise, length), dtype=dtype, device=torch.device("cuda", device)
                      if isinstance(device, int) else device)
    _check(lib().icm_noise(_dtype_code(dtype), int(seed), num_noise, length,
                           C.c_void_p(out.data_ptr()), _stream_ptr(out.device)), "icm_noise")
    return out


def sift_pass(h):
    """One sift pass over the rows of h ([I, N] device tensor).
    Returns (h_new, knot counts [I, 2] = (n_max, n_min) of h)."""
    torch = _torch()
    _require_cuda(h)
    if h.dim() == 1:
        h = h.unsqueeze(0)
    I, N = h.shape
    hn = torch.empty_like(h)
    nk = torch.empty((I, 2), dtype=torch.int32, device=h.device)
    _check(lib().icm_sift_pass(_dtype_code(h.dtype), C.c_void_p(h.data_ptr()), N, I,
                               C.c_void_p(hn.data_ptr()), C.c_void_p(nk.data_ptr()),
                               _stream_ptr(h.device)), "icm_sift_pass")
    return hn, nk


def sift_pass_static(h, out=None, nk=None):
    """sift_pass without any device allocation inside the library (its
    scratch is a fixed per-GPU array): I * N <= 2^21, I <= 4096.  `out` and
    `nk` may be given to avoid torch allocations too."""
    torch = _torch()
    _require_cuda(h)
    if h.dim() == 1:
        h = h.unsqueeze(0)
    I, N = h.shape
    hn = torch.empty_like(h) if out is None else out
    k = torch.empty((I, 2), dtype=torch.int32, device=h.device) if nk is None else nk
    _require_cuda(hn, k)
    if hn.shape != h.shape or hn.dtype != h.dtype or k.numel() != 2 * I or k.dtype != torch.int32:
        raise ValueError("out / nk do not match h")
    _check(lib().icm_sift_pass_static(_dtype_code(h.dtype), C.c_void_p(h.data_ptr()), N, I,
                                      C.c_void_p(hn.data_ptr()), C.c_void_p(k.data_ptr()),
                                      _stream_ptr(h.device)), "icm_sift_pass_static")
    return hn, k


def iceemdan(x, num_noise, num_imfs, max_sifts=10, stop_mode=STOP_FIXED, noise_strength=0.2,
             seed=1234, noise=None, noise_max_sifts=None):
    """Decompose one channel x (1-D device tensor) -> [num_imfs, N]."""
    plan = Plan([x.numel()], num_noise, num_imfs, dtype=x.dtype, stop_mode=stop_mode,
                max_sifts=max_sifts, noise_max_sifts=noise_max_sifts,
                noise_strength=noise_strength, seed=seed, use_graph=False, device=x.device)
    try:
        return plan.forward([x.contiguous()], noise=noise)[0]
    finally:
        plan.close()


def _autograd_function():
    torch = _torch()

    class ICEEMDANFunction(torch.autograd.Function):
        """x (1-D device tensor) -> [num_imfs, N] with a HIP backward.

        The extrema sets (and endpoint-fix branches) of every sift pass are
        held at the forward's values, so the backward is the exact gradient
        of the piecewise-linear map the forward evaluated."""

        @staticmethod
        def forward(ctx, x, plan, noise):
            out = plan.forward([x.detach().contiguous()], noise=noise)[0]
            ctx.plan = plan
            ctx.count = plan.forward_count
            return out

        @staticmethod
        def backward(ctx, g):
            if ctx.plan.forward_count != ctx.count:
                raise RuntimeError("the plan ran another forward before this backward")
            return ctx.plan.backward([g.contiguous()])[0], None, None

    return ICEEMDANFunction


_FN = None


def iceemdan_autograd(x, plan, noise=None):
    """Differentiable decomposition of one channel with a FIXED-mode plan
    whose single channel has x's length (torch.autograd)."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(x, plan, noise)


# ---- mirror of the reference entry points (cudaICEEMDAN.cuh:25-26) ----

def iceemdanS(numNoise, SignalLength, num_IMFs, max_iter, d_x, d_y, d_IMFs, noiseStrength):
    """Same arguments as the reference's iceemdanS; d_x (int32), d_y, d_IMFs
    (float32) are device tensors.  Returns the CEEMDAN-phase seconds, or a
    negative value on error."""
    torch = _torch()
    for t, dt in ((d_x, torch.int32), (d_y, torch.float32), (d_IMFs, torch.float32)):
        _require_cuda(t)
        if t.dtype != dt:
            raise TypeError("iceemdanS expects int32 d_x and float32 d_y / d_IMFs")
    return lib().icm_iceemdanS(numNoise, SignalLength, num_IMFs, max_iter,
                               C.c_void_p(d_x.data_ptr()), C.c_void_p(d_y.data_ptr()),
                               C.c_void_p(d_IMFs.data_ptr()), float(noiseStrength))


def iceemdanD(numNoise, SignalLength, num_IMFs, max_iter, d_x, d_y, d_IMFs, noiseStrength):
    """Double-precision twin of iceemdanS (cudaICEEMDAN.cu:1340)."""
    torch = _torch()
    for t, dt in ((d_x, torch.int32), (d_y, torch.float64), (d_IMFs, torch.float64)):
        _require_cuda(t)
        if t.dtype != dt:
            raise TypeError("iceemdanD expects int32 d_x and float64 d_y / d_IMFs")
    return lib().icm_iceemdanD(numNoise, SignalLength, num_IMFs, max_iter,
                               C.c_void_p(d_x.data_ptr()), C.c_void_p(d_y.data_ptr()),
                               C.c_void_p(d_IMFs.data_ptr()), float(noiseStrength))


# ---- statistics kernels (statistics.cuh:25-44) ----

def _stats_call(fn, dtype, *args):
    st = fn(_dtype_code(dtype), *args)
    if st != 0:
        raise RuntimeError(f"statistics kernel launch failed ({st})")


def meanS(inp, N, channels, chan_means):
    """chan_means[ch] += sum(inp[ch*N:(ch+1)*N]) (statistics.cu:58-94)."""
    _require_cuda(inp, chan_means)
    _stats_call(lib().icm_stats_mean, inp.dtype, C.c_void_p(inp.data_ptr()), int(N),
                int(channels), C.c_void_p(chan_means.data_ptr()), _stream_ptr(inp.device))


def varianceS(inp, N, channels, chan_means, chan_vars):
    """chan_vars[ch] += sum((x - chan_means[ch])^2) (statistics.cu:96-131)."""
    _require_cuda(inp, chan_means, chan_vars)
    _stats_call(lib().icm_stats_variance, inp.dtype, C.c_void_p(inp.data_ptr()), int(N),
                int(channels), C.c_void_p(chan_means.data_ptr()),
                C.c_void_p(chan_vars.data_ptr()), _stream_ptr(inp.device))


def multiplyS(inp, out, alpha, N):
    """out[i] = inp[i] * alpha (statistics.cu:50-56)."""
    _require_cuda(inp, out)
    _stats_call(lib().icm_stats_multiply, inp.dtype, C.c_void_p(inp.data_ptr()),
                C.c_void_p(out.data_ptr()), float(alpha), int(N), _stream_ptr(inp.device))


meanD, varianceD, multiplyD = meanS, varianceS, multiplyS
