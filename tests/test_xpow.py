"""xpow (csrc/tensor.h), the power law's |rss/tau|^(n-1) for a non-integer exponent n, against
the C library's pow, which is what the reference's Fortran `**` evaluates
(various_functions.f90 solve_res / slip rates).  xpow is __host__ __device__, so it is checked
here in a host build: random x over 23 e-folds and exponents 6-14 (the n - 1 of n in 7-15), the
relative error bounded by the exp(y ln x) form's |y ln x| eps growth, plus the special values."""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = r"""
#include "tensor.h"
#include <cmath>
#include <cstdio>
#include <random>
int main() {
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> lx(-20.0, 3.0), yy(6.0, 14.0);
    double worst = 0.0, worst_scaled = 0.0;
    for (long i = 0; i < 4000000; i++) {
        const double x = std::exp(lx(g)), y = yy(g);
        const double a = cg::xpow(x, y), b = std::pow(x, y);
        if (b == 0.0) { if (a != 0.0) { printf("FAIL zero %g %g\n", x, y); return 1; } continue; }
        const double rel = std::fabs(a - b) / b;
        const double t = std::fabs(y * std::log(x));
        if (rel > worst) worst = rel;
        // error budget of exp(y ln x): a few ulp plus |y ln x| ulp of the exponent
        const double scaled = rel / ((4.0 + t) * 2.220446049250313e-16);
        if (scaled > worst_scaled) worst_scaled = scaled;
    }
    printf("%.6g %.6g\n", worst, worst_scaled);
    printf("%.17g %.17g %.17g %.17g\n", cg::xpow(0.0, 9.5), cg::xpow(1.0, 9.5), cg::xpow(1e-300, 9.5),
           cg::xpow(0.5, 1.0));
    printf("%.17g %.17g\n", cg::xpow(2.0, 10.0), cg::xpow(3.0, 0.5));
    return 0;
}
"""


def test_xpow_matches_libm_pow(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    csrc = Path(__file__).resolve().parents[1] / "creep_modeling_amd" / "csrc"
    (tmp_path / "t.cpp").write_text(SRC)
    exe = tmp_path / "t"
    r = subprocess.run([hipcc, "-O2", "-std=c++17", "-I", str(csrc), str(tmp_path / "t.cpp"), "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout
    lines = out.stdout.split("\n")
    worst, worst_scaled = map(float, lines[0].split())
    assert worst < 1e-13, worst  # |y ln x| reaches 280 in this sample (x = e^-20, y = 14)
    assert worst_scaled < 4.0, worst_scaled
    z, one, tiny, half = map(float, lines[1].split())
    assert z == 0.0 and one == 1.0 and tiny == 0.0
    assert abs(half - 0.5) <= 2.3e-16
    p, s = map(float, lines[2].split())
    assert abs(p - 1024.0) <= 4 * 2.3e-16 * 1024.0
    assert abs(s - 3.0 ** 0.5) <= 4 * 2.3e-16 * 1.8
