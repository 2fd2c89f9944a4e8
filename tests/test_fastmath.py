"""
The device sincos / reciprocal helpers (ionization_amd/csrc/fastmath.h) compiled for the host
with g++ and compared with long-double libm: sin / cos within 2 ulp over |x| <= 1000, and the
small-angle form within 2 ulp over |x| <= 1/8.
"""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include "fastmath.h"
#include <cmath>
#include <cstdio>
#include <random>
static double ulp_err(double got, long double ref) {
    double r = (double)ref;
    double u = std::nextafter(std::fabs(r), INFINITY) - std::fabs(r);
    return (double)(fabsl((long double)got - ref) / u);
}
int main() {
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const double scales[4] = {1000.0, 20.0, 1.0, 1e-3};
    double ws = 0, wc = 0;
    for (int i = 0; i < 400000; ++i) {
        double x = U(rng) * scales[i % 4];
        double s, c;
        tdse_sincos(x, &s, &c);
        ws = std::fmax(ws, ulp_err(s, sinl((long double)x)));
        wc = std::fmax(wc, ulp_err(c, cosl((long double)x)));
    }
    // the small-angle form on its domain |x| <= 1/8
    double ss = 0, sc = 0;
    for (int i = 0; i < 400000; ++i) {
        double x = U(rng) * ((i % 3 == 0) ? kSmallAngle : (i % 3 == 1) ? 1e-2 : 1e-6);
        double s, c;
        tdse_sincos_small(x, &s, &c);
        ss = std::fmax(ss, ulp_err(s, sinl((long double)x)));
        sc = std::fmax(sc, ulp_err(c, cosl((long double)x)));
    }
    // exact special values used by the kernels: a rotation by 0 is the identity
    double s0, c0, t0, u0;
    tdse_sincos(0.0, &s0, &c0);
    tdse_sincos_small(0.0, &t0, &u0);
    printf("%.6f %.6f %d %.6f %.6f\n", ws, wc, (s0 == 0.0 && c0 == 1.0 && t0 == 0.0 && u0 == 1.0) ? 1 : 0, ss, sc);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_sincos_accuracy(tmp_path):
    src = tmp_path / "fm.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "fm"
    inc = os.path.join(ROOT, "ionization_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{inc}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    ws, wc, ident, ss, sc = float(out[0]), float(out[1]), int(out[2]), float(out[3]), float(out[4])
    assert ws <= 2.0 and wc <= 2.0, (ws, wc)
    assert ss <= 2.0 and sc <= 2.0, (ss, sc)
    assert ident == 1
