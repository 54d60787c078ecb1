// Device scalar arithmetic of the hot path: complex helpers, sin/cos and the
// Helmholtz point kernel. Translation units that must reproduce pivots bit for
// bit are compiled with -ffp-contract=off; every fused multiply-add below is
// written as an explicit fma().
#pragma once

#include <hip/hip_runtime.h>

namespace nca {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double cabs1(double2 a) { return fabs(a.x) + fabs(a.y); }
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    const double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
// acc + a * b with explicit fused operations (matvec accumulation).
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 acc) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
    return acc;
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// pi/2 in three parts (HI = RN(pi/2), MID = RN(pi/2 - HI), LO = RN(pi/2 - HI - MID)),
// 2/pi and 4 pi rounded to double.
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr double kPio2Hi = 0x1.921fb54442d18p+0;
constexpr double kPio2Mid = 0x1.1a62633145c07p-54;
constexpr double kPio2Lo = -0x1.f1976b7ed8fbcp-110;
constexpr double kFourPi = 0x1.921fb54442d18p+3;

// sin and cos of x: Cody-Waite reduction y = x - q pi/2 (the first fma step is
// exact for |q| < 2^20), Taylor polynomials of sin through y^19 and of cos
// through y^20 on |y| <= pi/4 with coefficients (-1)^k/(2k+1)! and
// (-1)^k/(2k)!, quadrant from q mod 4.
__device__ __forceinline__ void sincos_cw(double x, double& s_out, double& c_out) {
    const double q = rint(x * kTwoOverPi);
    double y = fma(-q, kPio2Hi, x);
    y = fma(-q, kPio2Mid, y);
    y = fma(-q, kPio2Lo, y);
    const double z = y * y;
    double ps = -0x1.2f49b46814157p-57;
    ps = fma(ps, z, 0x1.952c77030ad4ap-49);
    ps = fma(ps, z, -0x1.ae7f3e733b81fp-41);
    ps = fma(ps, z, 0x1.6124613a86d09p-33);
    ps = fma(ps, z, -0x1.ae64567f544e4p-26);
    ps = fma(ps, z, 0x1.71de3a556c734p-19);
    ps = fma(ps, z, -0x1.a01a01a01a01ap-13);
    ps = fma(ps, z, 0x1.1111111111111p-7);
    ps = fma(ps, z, -0x1.5555555555555p-3);
    const double sn = fma(y * z, ps, y);
    double pc = 0x1.e542ba4020225p-62;
    pc = fma(pc, z, -0x1.6827863b97d97p-53);
    pc = fma(pc, z, 0x1.ae7f3e733b81fp-45);
    pc = fma(pc, z, -0x1.93974a8c07c9dp-37);
    pc = fma(pc, z, 0x1.1eed8eff8d898p-29);
    pc = fma(pc, z, -0x1.27e4fb7789f5cp-22);
    pc = fma(pc, z, 0x1.a01a01a01a01ap-16);
    pc = fma(pc, z, -0x1.6c16c16c16c17p-10);
    pc = fma(pc, z, 0x1.5555555555555p-5);
    const double cs = fma(z * z, pc, fma(-0.5, z, 1.0));
    const long long n = (long long)q;
    switch (n & 3) {
        case 0: s_out = sn; c_out = cs; break;
        case 1: s_out = cs; c_out = -sn; break;
        case 2: s_out = -sn; c_out = -cs; break;
        default: s_out = -cs; c_out = sn; break;
    }
}

// G(x, y) = cis(sign k r) / (4 pi r), r = |x - y|, zero at r == 0.
__device__ __forceinline__ double2 helmholtz(double k, int sign, double x0, double x1, double x2, double y0,
                                             double y1, double y2) {
    const double dx = x0 - y0, dy = x1 - y1, dz = x2 - y2;
    double r2 = dx * dx + dy * dy;
    r2 = r2 + dz * dz;
    if (r2 == 0.0) return make_double2(0.0, 0.0);
    const double r = sqrt(r2);
    double s, c;
    sincos_cw(k * r, s, c);
    const double d = kFourPi * r;
    double2 g = make_double2(c / d, s / d);
    if (sign < 0) g.y = -g.y;
    return g;
}

}  // namespace nca
