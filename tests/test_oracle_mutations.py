"""The oracle's stage-algebra pins catch plausible slips.

Each case plants one mistake in a copy of oracle/iceemdan_oracle.c, builds
it into a temporary library, and runs the stage-algebra tests of
test_oracle.py against it in a child process: at least one of them must
fail.  The committed fixtures are left out, since they would catch every
change; the pins have to stand without them.

Planted slips (reference lines whose meaning each pin holds):
  E_{k+2} in place of E_{k+1} at stage k        cudaICEEMDAN.cu:80-90, :1149
  stage 0 without the 1/std(E_1^i) factor        cudaICEEMDAN.cu:68-78
  E_{k+1} normalised at stages k >= 1            cudaICEEMDAN.cu:80-90
  stage 0 reading E_2                            cudaICEEMDAN.cu:1011-1024
  h - m exported when the Rilling rule stops     cudaICEEMDAN.cu:527-544
"""
import os
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "iceemdan_oracle.c")

MUTANTS = {
    "noise_mode_index": (
        "const double* Ek = modes + (long)k * IN;",
        "const double* Ek = modes + (long)(k + 1 < M - 1 ? k + 1 : k) * IN;"),
    "stage0_unnormalised": (
        "x_i[p] = xt[p] + eps0 * E[p] * inv_i[i];",
        "x_i[p] = xt[p] + eps0 * E[p];"),
    "stagek_normalised": (
        "x_i[p] = r[p] + eps0 * E[p] * sq;",
        "x_i[p] = r[p] + eps0 * E[p] * sq * inv_i[i];"),
    "stage0_reads_E2": (
        "x_i[p] = xt[p] + eps0 * E[p] * inv_i[i];",
        "x_i[p] = xt[p] + eps0 * E[p + IN] * inv_i[i];"),
    "rilling_exports_h_minus_m": (
        "            if (!cont || j == S_or_cap - 1) break;\n"
        "            double* t = h; h = hn; hn = t;",
        "            double* t = h; h = hn; hn = t;\n"
        "            if (!cont || j == S_or_cap - 1) break;"),
}

CHILD = textwrap.dedent("""
    import sys
    sys.path.insert(0, {root!r})
    sys.path.insert(0, {tests!r})
    from oracle import oracle as orc
    orc.LIB_PATH = {so!r}
    orc._lib = None
    orc.lib()
    import test_oracle as t
    checks = [
        lambda: t.test_noise_modes_of_exact_alternation_and_ramp(orc),
        lambda: t.test_stage_noise_normalisation_and_index(orc),
        lambda: t.test_rilling_cap_one_exports_unsifted(orc, 1),
        lambda: t.test_rilling_cap_one_exports_unsifted(orc, 2000),
        lambda: t.test_iceemdan_completeness(orc),
        lambda: t.test_iceemdan_eps0_zero_is_noise_free(orc),
    ]
    failed = 0
    for f in checks:
        try:
            f()
        except AssertionError:
            failed += 1
    print("FAILED_PINS", failed)
""")


def _run_pins(so):
    code = CHILD.format(root=ROOT, tests=os.path.join(ROOT, "tests"), so=so)
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-2000:]
    line = [l for l in p.stdout.splitlines() if l.startswith("FAILED_PINS")][-1]
    return int(line.split()[1])


def _build(src_text, tmp_path, name):
    c = tmp_path / f"{name}.c"
    so = tmp_path / f"{name}.so"
    c.write_text(src_text)
    from oracle.build_oracle import CFLAGS
    subprocess.run(["gcc", *CFLAGS, "-I", os.path.join(ROOT, "oracle"), str(c), "-o", str(so), "-lm"],
                   check=True)
    return str(so)


def test_unmutated_oracle_passes_the_pins(tmp_path):
    so = _build(open(SRC).read(), tmp_path, "plain")
    assert _run_pins(so) == 0


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_fails_a_pin(tmp_path, name):
    text = open(SRC).read()
    old, new = MUTANTS[name]
    assert text.count(old) == 1, name
    so = _build(text.replace(old, new), tmp_path, name)
    assert _run_pins(so) >= 1, name
