// Double-precision helpers for the mesh kernels, written from the mathematics:
//  * tdse_sincos: Cody-Waite reduction by pi/2 with the three-part split of
//    pi/2 published in fdlibm (33 + 33 + 53 bits, so k * part is exact for
//    |k| < 2^20), then the Taylor series of sin (to r^17) and cos (to r^18) on
//    |r| <= pi/4, where the first omitted terms are below 1e-19. Valid for
//    |x| < 2^20 pi/2; the angles of the split-operator rotations are a few
//    radians at most.
//  * tdse_sincos_small: the same series without reduction for |x| <= 1/8.
//  * tdse_rcp: hardware reciprocal estimate refined to third order.
// The header also compiles for the host (without HIP) so the same code can
// be checked against long-double libm on the CPU.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define TDSE_HD __host__ __device__ __forceinline__
#else
#define TDSE_HD inline
#endif

namespace tdse_fm {
constexpr double kTwoOverPi = 0.6366197723675814;
constexpr double kPio2_1 = 1.5707963267341256;     // 0x1.921fb544p+0
constexpr double kPio2_2 = 6.077100506303966e-11;  // 0x1.0b4611a6p-34
constexpr double kPio2_3 = 2.0222662487959506e-21;
}  // namespace tdse_fm

TDSE_HD void tdse_sincos(double x, double* s, double* c) {
    using namespace tdse_fm;
    const double k = std::rint(x * kTwoOverPi);
    double r = std::fma(-k, kPio2_1, x);
    r = std::fma(-k, kPio2_2, r);
    r = std::fma(-k, kPio2_3, r);
    const double r2 = r * r;
    // sin r = r + r^3 (-1/3! + r^2 (1/5! - ...)), 1/(2i+1)! for i up to 8
    double ps = 2.8114572543455206e-15;
    ps = std::fma(ps, r2, -7.647163731819816e-13);
    ps = std::fma(ps, r2, 1.6059043836821613e-10);
    ps = std::fma(ps, r2, -2.505210838544172e-08);
    ps = std::fma(ps, r2, 2.7557319223985893e-06);
    ps = std::fma(ps, r2, -0.0001984126984126984);
    ps = std::fma(ps, r2, 0.008333333333333333);
    ps = std::fma(ps, r2, -0.16666666666666666);
    const double sr = std::fma(ps * r2, r, r);
    // cos r = 1 - r^2 / 2 + r^4 (1/4! - r^2 (1/6! - ...)), 1/(2i)! for i up to 9
    double pc = -1.5619206968586225e-16;
    pc = std::fma(pc, r2, 4.779477332387385e-14);
    pc = std::fma(pc, r2, -1.1470745597729725e-11);
    pc = std::fma(pc, r2, 2.08767569878681e-09);
    pc = std::fma(pc, r2, -2.755731922398589e-07);
    pc = std::fma(pc, r2, 2.48015873015873e-05);
    pc = std::fma(pc, r2, -0.001388888888888889);
    pc = std::fma(pc, r2, 0.041666666666666664);
    const double cr = std::fma(pc * r2, r2, std::fma(-0.5, r2, 1.0));
    // quadrant q = k mod 4: (sin, cos) = (s, c), (c, -s), (-s, -c), (-c, s)
    const int q = static_cast<int>(static_cast<long long>(k) & 3);
    double ss = (q & 1) ? cr : sr;
    double cc = (q & 1) ? sr : cr;
    if (q & 2) ss = -ss;
    if ((q + 1) & 2) cc = -cc;
    *s = ss;
    *c = cc;
}

// sin and cos for |x| <= 1/8 without reduction: the Taylor series of sin to
// x^11 and of cos to x^12, whose first omitted terms are below 3e-21 relative
// (about 14 operations against about 27 for tdse_sincos; the split-operator
// angles are this small away from the peak of a pulse).
constexpr double kSmallAngle = 0.125;
TDSE_HD void tdse_sincos_small(double x, double* s, double* c) {
    const double x2 = x * x;
    double ps = -2.505210838544172e-08;  // -1/11!
    ps = std::fma(ps, x2, 2.7557319223985893e-06);
    ps = std::fma(ps, x2, -0.0001984126984126984);
    ps = std::fma(ps, x2, 0.008333333333333333);
    ps = std::fma(ps, x2, -0.16666666666666666);
    *s = std::fma(ps * x2, x, x);
    double pc = 2.08767569878681e-09;  // 1/12!
    pc = std::fma(pc, x2, -2.755731922398589e-07);
    pc = std::fma(pc, x2, 2.48015873015873e-05);
    pc = std::fma(pc, x2, -0.001388888888888889);
    pc = std::fma(pc, x2, 0.041666666666666664);
    *c = std::fma(pc * x2, x2, std::fma(-0.5, x2, 1.0));
}

TDSE_HD double tdse_rcp(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    // 1/d = r0 / (1 - e) = r0 (1 + e + e^2 + ...), e = 1 - d r0: the estimate
    // (relative error below 2^-24 on gfx950) refined by the series to third
    // order, error ~e^3 below the final rounding; one FMA fewer than two
    // Newton steps
    const double r = __builtin_amdgcn_rcp(d);
    const double e = std::fma(-d, r, 1.0);
    return std::fma(r, std::fma(e, e, e), r);
#else
    return 1.0 / d;
#endif
}
