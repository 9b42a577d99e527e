"""C-ABI checks that need no GPU: the library loads, exports every entry
point declared in include/*.h, and validates arguments on the host."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared():
    """(name, is_c_linkage) of every function declared in include/*.h."""
    out = []
    for h in sorted(os.listdir(INC)):
        src = open(os.path.join(INC, h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        c_linkage = 'extern "C"' in src
        # prototypes ending in ';' (inline device helpers with bodies are skipped)
        for m in re.finditer(r"(?:__global__\s+)?\b(?:int|double|void|const\s+char\s*\*)\s+(\w+)\s*\([^;{]*\)\s*;", src):
            name = m.group(1)
            is_kernel = "__global__" in m.group(0)
            out.append((name, c_linkage and not is_kernel, h))
    return out


def exported(lib_path):
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    res = subprocess.run([nm, "-DC", "--defined-only", lib_path], capture_output=True, text=True,
                         check=True).stdout
    names = set()
    for line in res.splitlines():
        parts = line.split(None, 2)
        if len(parts) == 3:
            names.add(parts[2].split("(")[0])
    return names


def test_library_loads(icm):
    L = icm.lib()
    assert L.icm_status_string(0) == b"ok"
    assert L.icm_status_string(-5).startswith(b"sample-index")


def test_every_declared_symbol_is_exported(icm):
    decl = declared()
    names = {d[0] for d in decl}
    # the headers must declare the reference entry points and the C ABI
    for must in ("iceemdanS", "iceemdanD", "meanS", "varianceS", "multiplyS", "meanD",
                 "varianceD", "multiplyD", "icm_plan_create", "icm_forward", "icm_backward",
                 "icm_noise", "icm_sift_pass"):
        assert must in names, must
    syms = exported(icm.LIB_PATH)
    L = icm.lib()
    for name, c_linkage, header in decl:
        if c_linkage:
            assert hasattr(L, name), f"{name} ({header}) not exported"
        else:
            assert name in syms, f"{name} ({header}) not exported"


def test_python_signatures_cover_c_abi(icm):
    c_names = {d[0] for d in declared() if d[1]}
    missing = c_names - set(icm._SIGS)
    assert not missing, missing


def _cfg(icm, **kw):
    base = dict(lengths=[4096], num_noise=10, num_imfs=5)
    base.update(kw)
    return icm.validate_config(**base)


def test_host_validation(icm):
    assert _cfg(icm) == 0
    bad = [dict(lengths=[3]), dict(num_noise=0), dict(num_imfs=1),
           dict(max_sifts=0), dict(stop_mode=7), dict(dtype_code=5),
           dict(noise_strength=float("nan")), dict(lengths=[20_000_000]), dict(boundary=2),
           dict(num_noise=70000),
           # one epoch per pass: a stage's passes must stay below the 65536
           # control words, so caps above 32000 are refused
           dict(max_sifts=32001), dict(noise_max_sifts=32001),
           dict(stop_mode=1, max_sifts=65535),
           # per-channel launches put the channel on gridDim.y
           dict(lengths=[16] * 65536)]
    for b in bad:
        assert _cfg(icm, **b) == -1, b
    for ok in (dict(max_sifts=32000), dict(stop_mode=1, noise_max_sifts=32000),
               dict(lengths=[16] * 65535, num_noise=1)):
        assert _cfg(icm, **ok) == 0, ok


def test_full_128_channel_job_fits_one_plan(icm):
    """The whole 128 x 244032, I = 500 forward is one plan (61M tiles, 1.56e10
    samples per launch: grid-stride and work-queue kernels, no per-launch
    work-item limit); it needs about 200 GB in float32, within one MI355X's
    288 GB.  Phases of more than 2^28 tiles are refused (32-bit tile
    indices)."""
    full = dict(lengths=[244032] * 128, num_noise=500, num_imfs=16, dtype_code=0)
    assert icm.validate_config(**full) == 0
    b = icm.plan_bytes(**full)
    assert 150e9 < b < 288e9, b
    assert icm.validate_config(lengths=[244032] * 1200, num_noise=1000, num_imfs=4) == -1


def test_tuning_candidates(icm):
    """Candidate c: segments of 4 << (c / 2) tiles, small tables when c is
    odd; -1 is the default geometry."""
    seen = set()
    for c in range(icm.NUM_TUNING_CANDIDATES):
        seg, small = icm.tuning_candidate(c)
        assert seg == 4 << (c // 2) and small == (c & 1)
        seen.add((seg, small))
    assert len(seen) == icm.NUM_TUNING_CANDIDATES
    assert icm.tuning_candidate(-1) == (0, -1)
    with pytest.raises(RuntimeError):
        icm.tuning_candidate(icm.NUM_TUNING_CANDIDATES)


def test_plan_bytes_estimate(icm):
    b = icm.plan_bytes(lengths=[244032], num_noise=500, num_imfs=16)
    # h ping-pong + 15 noise modes + residual, float64: about 18 GB plus knots
    assert 17e9 < b < 30e9
    b32 = icm.plan_bytes(lengths=[244032], num_noise=500, num_imfs=16, dtype_code=0)
    assert b32 < b


def test_product_path_has_no_cpu_fallback(icm):
    """The package raises instead of computing on the CPU when no GPU tensor
    is given."""
    import torch
    x = torch.zeros(100, dtype=torch.float64)
    with pytest.raises((ValueError, RuntimeError)):
        icm.sift_pass(x.reshape(1, -1))


def test_plan_above_memory_budget_is_refused_before_allocating(icm, monkeypatch):
    """ICM_MEMORY_BUDGET: a plan whose size estimate exceeds it is refused with
    ICM_ERR_OUT_OF_MEMORY and a message, before any device call (this runs
    without a GPU)."""
    monkeypatch.setenv("ICM_MEMORY_BUDGET", str(10**9))
    cfg, _keep = icm.make_config([244032], 500, 16)
    h = C.c_void_p()
    assert icm.lib().icm_plan_create(C.byref(cfg), C.byref(h)) == -2
    assert not h.value
    assert b"budget" in icm.lib().icm_last_error()
    assert "budget" in str(icm.ICMError(-2))
