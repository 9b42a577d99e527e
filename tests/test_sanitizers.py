"""Sanitizer builds on the CPU.

* The FP64 oracle (oracle/iceemdan_oracle.c) and a driver of every oracle
  entry point (tests/sanitize/oracle_driver.c) built with
  -fsanitize=address,undefined: the driver must exit 0 with no report.
* The library's host code (every .hip under csrc/, host side only, with
  -fsanitize=address,undefined on the host pass) and a driver of its
  host-only entry points (argument validation, plan-size estimates for
  every benchmark shape, tuning-candidate decoding, status strings, the
  memory-budget refusal): likewise.
GPU code is not sanitized (no GPU address sanitizer here)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def test_oracle_under_address_and_undefined_sanitizers(tmp_path):
    exe = tmp_path / "oracle_driver"
    subprocess.run(["gcc", "-O1", "-g", "-std=gnu11", "-fno-omit-frame-pointer", "-ffp-contract=off",
                    "-mfma", "-fopenmp", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-I", os.path.join(ROOT, "oracle"),
                    os.path.join(ROOT, "oracle", "iceemdan_oracle.c"),
                    os.path.join(ROOT, "tests", "sanitize", "oracle_driver.c"),
                    "-o", str(exe), "-lm"], check=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, env=ENV, timeout=600)
    assert p.returncode == 0, p.stderr[-4000:]
    assert "oracle driver ok" in p.stdout
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr


HOST_DRIVER = r"""
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "iceemdan_mi355x.h"

static int check(icm_config* c, int want)
{
    int64_t b = 0;
    int v = icm_validate_config(c);
    if (v != want) { printf("validate %d want %d\n", v, want); return 1; }
    if (v == 0 && (icm_plan_bytes(c, &b) != 0 || b <= 0)) { printf("bytes\n"); return 1; }
    return 0;
}

int main(void)
{
    int bad = 0;
    static int64_t lens[128];
    icm_config c;
    icm_config_default(&c);
    for (int i = 0; i < 128; ++i) lens[i] = 8192 + 1937 * i;
    c.lengths = lens;
    int shapes[][4] = {{1, 500, 16, 1}, {64, 200, 14, 0}, {32, 100, 12, 1}, {128, 500, 16, 0},
                       {3, 1, 2, 1}};
    for (int k = 0; k < 5; ++k) {
        c.num_channels = shapes[k][0];
        c.num_noise = shapes[k][1];
        c.num_imfs = shapes[k][2];
        c.dtype = shapes[k][3];
        c.keep_records = k & 1;
        bad |= check(&c, 0);
    }
    c.num_imfs = 1;
    bad |= check(&c, -1);
    c.num_imfs = 4;
    c.lengths = NULL;
    bad |= check(&c, -1);
    for (int cand = -2; cand <= 8; ++cand) {
        int seg = -9, small = -9;
        int r = icm_tuning_candidate(cand, &seg, &small);
        if ((cand < -1 || cand > 7) != (r != 0)) { printf("cand %d\n", cand); bad = 1; }
    }
    for (int s = -7; s <= 0; ++s)
        if (!icm_status_string(s) || !strlen(icm_status_string(s))) bad = 1;
    /* a plan above the memory budget is refused before any device call */
    setenv("ICM_MEMORY_BUDGET", "1000000", 1);
    int64_t one = 244032;
    c.num_channels = 1;
    c.lengths = &one;
    c.num_noise = 500;
    c.num_imfs = 16;
    c.dtype = 1;
    icm_plan p = NULL;
    if (icm_plan_create(&c, &p) != ICM_ERR_OUT_OF_MEMORY || p) { printf("budget\n"); bad = 1; }
    if (!strstr(icm_last_error(), "budget")) { printf("msg %s\n", icm_last_error()); bad = 1; }
    printf(bad ? "host driver FAILED\n" : "host driver ok\n");
    return bad;
}
"""


def test_library_host_code_under_address_and_undefined_sanitizers(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    csrc = os.path.join(ROOT, "ceemdan-gpu_amd", "csrc")
    srcs = sorted(os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith(".hip"))
    lib = tmp_path / "libicm_host_asan.so"
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-omit-frame-pointer"]
    from concurrent.futures import ThreadPoolExecutor

    def compile_one(src):
        obj = tmp_path / (os.path.basename(src) + ".o")
        subprocess.run([hipcc, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC",
                        "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", *san, "-c", src, "-o", str(obj)],
                       check=True, capture_output=True)
        return str(obj)

    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    subprocess.run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *san, *objs,
                    "-o", str(lib)], check=True, capture_output=True)
    drv = tmp_path / "host_driver.cpp"
    drv.write_text(HOST_DRIVER)
    exe = tmp_path / "host_driver"
    # the driver is host code only: the same clang as hipcc (one sanitizer runtime)
    clangxx = os.path.join(os.path.dirname(os.path.realpath(hipcc)), "..", "llvm", "bin", "clang++")
    if not os.path.exists(clangxx):
        clangxx = "/opt/rocm/llvm/bin/clang++"
    subprocess.run([clangxx, "-O1", "-g", "-fsanitize=address,undefined",
                    "-I", os.path.join(ROOT, "include"), str(drv), str(lib),
                    f"-Wl,-rpath,{tmp_path}", "-o", str(exe)], check=True, capture_output=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, env=ENV, timeout=300)
    assert p.returncode == 0, (p.stdout[-2000:], p.stderr[-4000:])
    assert "host driver ok" in p.stdout
    assert "runtime error" not in p.stderr
