"""The fast mode's logarithm (ln_pos, csrc/ppes_math.hpp) against the long
double logarithm of the host C library: within 1.5 ulp on random arguments
over the whole positive range (subnormals included), near 1 where ln x is
small, and at the reduction boundary 1/sqrt(2); the library's values for 0,
negative, infinite and NaN arguments. The header runs the same IEEE
operations on the host as on the device (frexp, correctly rounded division,
fused multiply-add), so the host build checks the device's arithmetic. A
planted mistake (one series coefficient off by 1e-6 relative) must fail."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pp-es-dg-mhd_amd", "csrc")

PROGRAM = r"""
#include "ppes_math.hpp"
#include <cmath>
#include <cstdio>
#include <random>

static double ulps(double got, long double ref)
{
    const double r = (double)ref;
    const double u = std::nextafter(std::fabs(r), INFINITY) - std::fabs(r);
    return (double)(std::fabs((long double)got - ref) / u);
}

int main()
{
    std::mt19937_64 gen(7);
    std::uniform_real_distribution<double> mant(0.5, 1.0), expo(-1074.0, 1024.0), near(0.25, 4.0);
    double worst = 0.0;
    auto check = [&](double x) {
        const double e = ulps(ppes::ln_pos(x), logl((long double)x));
        if (e > worst) worst = e;
    };
    for (int i = 0; i < 1000000; ++i) check(std::ldexp(mant(gen), (int)std::floor(expo(gen))));
    for (int i = 0; i < 1000000; ++i) check(near(gen));
    for (int i = 0; i < 200000; ++i) {
        const double d = std::ldexp(mant(gen), -(1 + (int)(52 * mant(gen))));
        check(1.0 + d);
        check(1.0 - d);
    }
    double b = 0.70710678118654752;
    for (int i = 0; i < 1000; ++i) { check(b); b = std::nextafter(b, 0.0); }
    b = 0.70710678118654752;
    for (int i = 0; i < 1000; ++i) { check(b); b = std::nextafter(b, 1.0); }
    check(4.9406564584124654e-324);
    check(1.7976931348623157e308);
    std::printf("%.6f %.17g %.17g %.17g %.17g %.17g\n", worst, ppes::ln_pos(1.0), ppes::ln_pos(0.0),
                ppes::ln_pos(-1.0), ppes::ln_pos(INFINITY), ppes::ln_pos(NAN));
    return 0;
}
"""


def run(tmp_path, header_dir):
    src = tmp_path / "ln_pos_check.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "ln_pos_check"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", str(header_dir), "-o", str(exe), str(src)])
    out = subprocess.check_output([str(exe)], text=True).split()
    return float(out[0]), [float(v) for v in out[1:]]


def test_ln_pos_within_1p5_ulp(tmp_path):
    worst, special = run(tmp_path, CSRC)
    print(f"ln_pos: worst {worst:.3f} ulp")
    assert worst <= 1.5
    one, zero, neg, inf, nan = special
    assert one == 0.0 and zero == float("-inf") and inf == float("inf")
    assert neg != neg and nan != nan


def test_planted_coefficient_error_fails(tmp_path):
    text = open(os.path.join(CSRC, "ppes_math.hpp")).read()
    planted = text.replace("p = fma(p, z, 2.0 / 5.0);", "p = fma(p, z, 2.0 / 5.0 * (1.0 + 1e-6));")
    assert planted != text
    d = tmp_path / "planted"
    d.mkdir()
    (d / "ppes_math.hpp").write_text(planted)
    worst, _ = run(tmp_path, d)
    assert worst > 1.5
