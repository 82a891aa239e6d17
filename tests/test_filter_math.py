"""The matched filter's float64 ln and atan2 (exmarg_amd/csrc/exm_fmath.h, used by the band transform's
second pass) against the C library, on the CPU: the header is compiled for the host with hipcc and
evaluated on 4e6 inputs spanning 1e-30 .. 1e30 in magnitude, points near the axes, on the octant
boundaries and near the diagonal, and on the special cases (signed zeros, the axes).  The filter rounds
both results to float32; the bar is that no float32-rounded result differs from the C library's."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

DRIVER = r"""
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include "exm_fmath.h"

static int64_t ulps(float a, float b) {
    int32_t x, y;
    std::memcpy(&x, &a, 4);
    std::memcpy(&y, &b, 4);
    return x > y ? (int64_t)x - y : (int64_t)y - x;
}

int main() {
    std::mt19937_64 g(7);
    std::normal_distribution<double> n(0.0, 1.0);
    std::uniform_real_distribution<double> u(-300.0, 300.0);
    long long tot = 0, dl = 0, da = 0;
    double rl = 0.0, ra = 0.0;
    for (int k = 0; k < 4000000; ++k) {
        const double sc = std::pow(10.0, u(g) / 10.0);
        double x = n(g) * sc, y = n(g) * sc;
        if (k % 7 == 0) y *= 1e-9;
        if (k % 11 == 0) x *= 1e-12;
        if (k % 13 == 0) y = x * (0.41421356237309504880 + 1e-8 * n(g));
        if (k % 17 == 0) y = x * (1.0 + 1e-9 * n(g));
        const double p0 = x * x + y * y, p = p0 > 1e-38 ? p0 : 1e-38;
        if (!std::isfinite(p)) continue;
        ++tot;
        dl += ulps((float)(0.5 * std::log(p)), (float)(0.5 * exm::ln_pos(p))) != 0;
        da += ulps((float)std::atan2(y, x), (float)exm::atan2_f(y, x)) != 0;
        rl = std::fmax(rl, std::fabs(exm::ln_pos(p) - std::log(p)) / std::fmax(std::fabs(std::log(p)), 1e-300));
        ra = std::fmax(ra, std::fabs(exm::atan2_f(y, x) - std::atan2(y, x)) /
                               std::fmax(std::fabs(std::atan2(y, x)), 1e-300));
    }
    const double sp[][2] = {{0, 0}, {-0.0, 0}, {0, -1}, {-0.0, -1}, {1, 0}, {-1, 0}, {0, 1}, {-0.0, 1},
                            {1e-300, 1e-300}, {1e300, -1e300}, {-3, -4}, {3, -4}, {2, 2}, {-2, -2}};
    int bad = 0;
    for (const auto& v : sp) {
        const double r = std::atan2(v[0], v[1]), a = exm::atan2_f(v[0], v[1]);
        bad += (float)r != (float)a || std::signbit(r) != std::signbit(a);
    }
    for (double v : {1e-38, 0.5, 1.0, 2.0, 0.70710678118654752, 1.4142135623730951, 3.7, 1e300})
        bad += (float)std::log(v) != (float)exm::ln_pos(v);
    std::printf("%lld %lld %lld %.3g %.3g %d\n", tot, dl, da, rl, ra, bad);
    return 0;
}
"""


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is not installed")
def test_filter_math_rounds_like_libm(tmp_path):
    src = tmp_path / "fmath_check.hip"
    src.write_text(DRIVER)
    exe = tmp_path / "fmath_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-I", os.path.join(ROOT, "exmarg_amd", "csrc"),
                    "-o", str(exe), str(src)], check=True, capture_output=True, timeout=600)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=600).stdout.split()
    tot, dl, da, rl, ra, bad = int(out[0]), int(out[1]), int(out[2]), float(out[3]), float(out[4]), int(out[5])
    assert tot > 3900000
    assert dl == 0 and da == 0, (dl, da)
    assert rl < 2e-15 and ra < 1e-15, (rl, ra)
    assert bad == 0
