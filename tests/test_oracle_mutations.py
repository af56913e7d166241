"""The oracle's geometry and splat pins catch plausible mistakes (CPU).

Each case copies oracle/fip_oracle.c, plants one mistake in the copy,
compiles it and runs the pin tests of tests/test_oracle_imaging.py against it
in a child process (FIP_ORACLE_LIB names the broken library).  The pins must
fail for every mutant and pass for an unmodified copy built the same way.
"""
import os
import re
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

PINS = ("test_point_source_final_peak_on_sky_position or test_splat_subpixel_bilinear_corners "
        "or test_splat_first_moments_are_linear or test_point_source_pre_image_peak_in_pca_frame "
        "or test_point_source_on_pixel_known_amplitude")

# (name, [(pattern, replacement, expected count)], pin tests that must fail)
MUTANTS = [
    ("swap_V01_V10",
     [(r"\(double\)vin\[3\] \* b", "(double)vin[@] * b", 1),
      (r"\(double\)vin\[1\] \* a", "(double)vin[3] * a", 1),
      (r"vin\[@\]", "vin[1]", 1)],
     "test_point_source_final_peak_on_sky_position"),
    ("drop_abs_V22_divisor",
     [(r"\) / a22 \+ h;", ") + h;", 2)],
     "test_point_source_final_peak_on_sky_position"),
    ("signed_V22",
     [(r"const double a22 = fabs\(\(double\)vin\[8\]\);", "const double a22 = (double)vin[8];", 1)],
     "test_point_source_final_peak_on_sky_position"),
    ("swap_LL_MM_weights",
     [(r"const double wl0 = 1\.0 - LL \+ floor\(LL\), wl1 = LL - floor\(LL\);",
       "const double wl0 = 1.0 - MM + floor(MM), wl1 = MM - floor(MM);", 1),
      (r"const double wm0 = 1\.0 - MM \+ floor\(MM\), wm1 = MM - floor\(MM\);",
       "const double wm0 = 1.0 - LL + floor(LL), wm1 = LL - floor(LL);", 1)],
     "test_splat_subpixel_bilinear_corners"),
    ("transposed_projection_B",
     [(r"bin\[2 \* n\] = ", "bin[2 * n + @] = ", 1),
      (r"bin\[2 \* n \+ 1\] = ", "bin[2 * n] = ", 1),
      (r"bin\[2 \* n \+ @\] = ", "bin[2 * n + 1] = ", 1)],
     "test_point_source_pre_image_peak_in_pca_frame"),
    ("centred_projection_B",
     [(r"V\[0\]\[2\] \+ m0;", "V[0][2];", 1),
      (r"V\[1\]\[2\] \+ m1;", "V[1][2];", 1)],
     "test_point_source_on_pixel_known_amplitude"),
]


def _build(tmp, src_text, name):
    d = os.path.join(tmp, name)
    os.makedirs(d)
    shutil.copy(os.path.join(ORACLE, "fip_oracle.h"), d)
    with open(os.path.join(d, "fip_oracle.c"), "w") as f:
        f.write(src_text)
    so = os.path.join(d, "libfip_oracle.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-fopenmp", "-shared", "-o", so,
                    os.path.join(d, "fip_oracle.c"), "-lm"], check=True)
    return so


def _run_pins(so, selector):
    env = dict(os.environ, FIP_ORACLE_LIB=so)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_imaging.py"), "-k", selector],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout[-2000:]


def _source():
    with open(os.path.join(ORACLE, "fip_oracle.c")) as f:
        return f.read()


def test_unmutated_copy_passes_the_pins(tmp_path):
    so = _build(str(tmp_path), _source(), "control")
    rc, out = _run_pins(so, PINS)
    assert rc == 0, out


@pytest.mark.parametrize("name,edits,selector", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_fails_the_pins(tmp_path, name, edits, selector):
    src = _source()
    for pat, rep, count in edits:
        src, n = re.subn(pat, rep, src)
        assert n == count, f"{name}: pattern {pat!r} matched {n} times, expected {count}"
    so = _build(str(tmp_path), src, name)
    rc, out = _run_pins(so, selector)
    assert rc != 0, f"mutant {name} passed the pins:\n{out}"
