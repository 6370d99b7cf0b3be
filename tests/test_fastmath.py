"""The kernels' FP64 cube root and cos(acos(x)/3) (FP32 seeds refined in FP64,
csrc/kernels/fastmath.hpp), compiled for the host from the same source, are
within about one unit in the last place of long-double references, as the
libm functions they replace in the closed-form cubic are (realCubicSolution,
tools.cpp:350-393)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fastmath_accuracy(tmp_path):
    exe = str(tmp_path / "fastmath_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", os.path.join(ROOT, "tests", "cpp",
                    "fastmath_check.cpp"), "-o", exe, "-lm"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120, check=True).stdout
    res = {l.split()[0]: [float(x) for x in l.split()[1:]] for l in out.splitlines()}
    n, err, err_lib = res["cbrt"]
    assert n == 2000000 and err <= 1.0, res
    n, err, err_lib = res["cos_third_acos"]
    # samples below x = -0.99 take the library composition on both sides
    assert n == 2000000 and err <= max(1.0, err_lib), res
