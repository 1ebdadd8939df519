"""The kinetics' exponential (exp_fin, eu3d_amd/csrc/exp_fin.hpp) against the C library's exp.

The header is host-callable; a small C++ harness built here with g++ evaluates it on 10^7 seeded
arguments over [-708, 709] and [-2, 2] (the range of the rate exponents, and the reduced range
around zero) and at the edges (overflow, underflow, denormal results, signed zero, NaN), and
compares with std::exp bit pattern by bit pattern. A missing term of the series, a wrong
coefficient or the loss of the low part of ln2 moves results by more than one ulp or makes the
exactly rounded fraction collapse."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HDR = os.path.join(os.path.dirname(HERE), "eu3d_amd", "csrc")

HARNESS = r"""
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "exp_fin.hpp"
static int64_t ord(double d) { int64_t i; std::memcpy(&i, &d, 8); return i < 0 ? INT64_MIN - i : i; }
int main() {
    uint64_t s = 88172645463325252ull;
    long long within1 = 0, exact = 0, worse = 0, n = 0;
    for (long long k = 0; k < 10000000LL; k++) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        const double u = (double)(s >> 11) / 9007199254740992.0;
        const double x = (k & 1) ? -708.0 + 1417.0 * u : -2.0 + 4.0 * u;
        const long long d = std::llabs((long long)(ord(eu3d::exp_fin(x)) - ord(std::exp(x))));
        n++;
        if (d == 0) exact++;
        if (d <= 1) within1++; else worse++;
    }
    std::printf("%lld %lld %lld %lld\n", n, exact, within1, worse);
    const double e[] = {0.0, -0.0, 1.0, -1.0, 709.7, 710.0, 800.0, -708.3, -745.0, -746.0, -800.0, 1e-300};
    for (double x : e) std::printf("%.17g %.17g %.17g\n", x, eu3d::exp_fin(x), std::exp(x));
    std::printf("nan %d\n", (int)std::isnan(eu3d::exp_fin(std::nan(""))));
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("expfin")
    src, exe = d / "h.cpp", d / "h"
    src.write_text(HARNESS)
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-I", HDR, "-o", str(exe), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True, timeout=120).stdout.splitlines()
    return out


def test_within_one_ulp_of_libm(harness):
    n, exact, within1, worse = (int(v) for v in harness[0].split())
    assert n == 10_000_000 and worse == 0 and within1 == n
    # the measured share of exactly rounded results is about 94 %
    assert exact > 0.9 * n


def test_edges(harness):
    for line in harness[1:-1]:
        x, a, b = (float(v) for v in line.split())
        if b == 0.0 or b == float("inf"):
            assert a == b, (x, a, b)
        else:
            assert abs(a - b) <= 2.3e-16 * b, (x, a, b)
    assert harness[-1] == "nan 1"
