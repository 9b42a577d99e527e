"""The command-line driver (the reference's sample_binary_file.cu /
sample_synthetic_signal.cu counterpart): usage on the CPU, results on the GPU
equal to the library called directly."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ceemdan-gpu_amd", "bin", "iceemdan_file")
GOLD = os.path.join(ROOT, "tests", "golden")


def test_driver_usage_without_gpu():
    if not os.path.exists(EXE):
        pytest.skip("driver not built")
    p = subprocess.run([EXE], capture_output=True, text=True)
    assert p.returncode == 2 and "usage" in p.stderr
    p = subprocess.run([EXE, "--noise", "0", "out.bin", "in.bin"], capture_output=True, text=True)
    assert p.returncode == 2


@pytest.mark.gpu
def test_driver_matches_library(icm, dev, tmp_path):
    import torch
    assert os.path.exists(EXE)
    x = np.fromfile(os.path.join(GOLD, "eegSampleDataCH4.f32"), dtype="<f4")
    a, b = x[:3000].copy(), x[5000:9100].copy()
    pa, pb = tmp_path / "a.bin", tmp_path / "b.bin"
    a.tofile(pa)
    b.tofile(pb)
    print("case: driver single channel fixed fp64", flush=True)
    out = tmp_path / "modes.bin"
    p = subprocess.run([EXE, "--noise", "8", "--imfs", "5", "--fixed", "10", "--double", str(out),
                        str(pa)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    got = np.fromfile(out, dtype="<f4").reshape(5, 3000)
    ref = icm.iceemdan(torch.tensor(a.astype(np.float64), device=dev), 8, 5, max_sifts=10)
    assert np.array_equal(got, ref.cpu().numpy().astype(np.float32))

    print("case: driver two channels batched fp32", flush=True)
    p = subprocess.run([EXE, "--noise", "6", "--imfs", "4", "--fixed", "8", str(tmp_path / "m"),
                        str(pa), str(pb)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    plan = icm.Plan([3000, 4100], 6, 4, dtype=torch.float32, max_sifts=8, use_graph=False,
                    device=dev)
    refs = plan.forward([torch.tensor(a, device=dev), torch.tensor(b, device=dev)])
    plan.close()
    for c, (n, r) in enumerate(zip((3000, 4100), refs)):
        got = np.fromfile(tmp_path / f"m.{c}.bin", dtype="<f4").reshape(4, n)
        assert np.array_equal(got, r.cpu().numpy()), c

    print("case: driver two-tone drop-in", flush=True)
    out = tmp_path / "tt.bin"
    p = subprocess.run([EXE, "--noise", "100", "--imfs", "2", "--max-iter", "100", "--two-tone",
                        str(out)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "Execution time" in p.stdout
    got = np.fromfile(out, dtype="<f4").reshape(2, 1000)
    n = np.arange(1000)
    pi = 3.1415926   # the value sample_synthetic_signal.cu:32 uses
    y = np.sin(2 * pi * 0.065 * n) + np.where((n >= 500) & (n <= 748),
                                              np.sin(2 * pi * 0.255 * (n - 500)), 0.0)
    d_x = torch.arange(1000, dtype=torch.int32, device=dev)
    d_y = torch.tensor(y.astype(np.float32), device=dev)
    d_imfs = torch.empty(2000, dtype=torch.float32, device=dev)
    assert icm.iceemdanS(100, 1000, 2, 100, d_x, d_y, d_imfs, 0.2) >= 0
    # libm sin may differ from numpy's in the last bit of the float64 input
    assert np.max(np.abs(got - d_imfs.cpu().numpy().reshape(2, 1000))) < 1e-4
