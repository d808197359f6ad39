// pt_libm.h — portable double-precision sin/cos/acos/atan2 for the render path.
//
// The reference renderer calls the unqualified C functions `sin`, `cos`,
// `acos` and `atan2` on float arguments (e.g. Math/OrthonormalBasis.cpp:39-41,
// Core/Material/Lambert.cpp:69-72, Core/Material/GGX.cpp:193-194,
// Core/Environment.cpp:38-39,231-232). Under libstdc++ these bind to the
// *double* C library functions, so every such call is "widen to double,
// evaluate in double, round to float at the assignment" — and products such as
// `sinTheta * cos(phi)` are formed in double before that rounding.
//
// To make the GPU kernels and the host agree bit for bit, both evaluate these
// functions with the routines below: plain IEEE double additions and
// multiplications in a fixed order (no FMA: every translation unit that
// includes this is compiled with -ffp-contract=off), Cody–Waite argument
// reduction for sin/cos, and truncated Taylor series whose coefficients are the
// correctly rounded values of (-1)^k/(2k+1)!, (-1)^k/(2k)!,
// (2n)!/(4^n (n!)^2 (2n+1)) and (-1)^n/(2n+1). Every routine is accurate to
// about one double ulp, so after rounding to float the result equals glibc's
// except when the exact value lies within ~2^-52 (relative) of a float rounding
// midpoint.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define PT_HD __host__ __device__ __forceinline__
#else
#define PT_HD static inline
#endif

// Round to the nearest integer (ties to even) for |x| < 2^51, by adding and
// subtracting 1.5*2^52. Exact on any IEEE double unit in round-to-nearest mode
// (the sum is not re-associated: nothing here is compiled with -ffast-math).
PT_HD double pt_rint_small(double x)
{
    const double kMagic = 6755399441055744.0;
    const double t = x + kMagic;
    return t - kMagic;
}

PT_HD double pt_fabs_d(double x) { return __builtin_fabs(x); }

// sin(r) for |r| <= pi/4: r + r*z*P(z), z = r^2, P from (-1)^k/(2k+1)!, k = 1..9.
PT_HD double pt_kernel_sin(double r)
{
    const double z = r * r;
    double p = -0x1.2f49b46814157p-57;
    p = p * z + 0x1.952c77030ad4ap-49;
    p = p * z + -0x1.ae7f3e733b81fp-41;
    p = p * z + 0x1.6124613a86d09p-33;
    p = p * z + -0x1.ae64567f544e4p-26;
    p = p * z + 0x1.71de3a556c734p-19;
    p = p * z + -0x1.a01a01a01a01ap-13;
    p = p * z + 0x1.1111111111111p-7;
    p = p * z + -0x1.5555555555555p-3;
    // sin(-0) = -0: r + (-0 * z) * p would round to +0
    return r == 0.0 ? r : r + (r * z) * p;
}

// cos(r) for |r| <= pi/4: 1 + z*Q(z), Q from (-1)^k/(2k)!, k = 1..10.
PT_HD double pt_kernel_cos(double r)
{
    const double z = r * r;
    double q = 0x1.e542ba4020225p-62;
    q = q * z + -0x1.6827863b97d97p-53;
    q = q * z + 0x1.ae7f3e733b81fp-45;
    q = q * z + -0x1.93974a8c07c9dp-37;
    q = q * z + 0x1.1eed8eff8d898p-29;
    q = q * z + -0x1.27e4fb7789f5cp-22;
    q = q * z + 0x1.a01a01a01a01ap-16;
    q = q * z + -0x1.6c16c16c16c17p-10;
    q = q * z + 0x1.5555555555555p-5;
    // 1 - z/2 + z^2*(...): keep the large term separate for accuracy.
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + (z * z) * q);
}

// Simultaneous sin and cos of a double that came from a float (|x| < 2^20).
PT_HD void pt_sincos(double x, double* s, double* c)
{
    if (!(x == x)) { *s = x; *c = x; return; }
    // Cody–Waite: pi/2 = P1 + P1t (+ ~3.5e-27), P1 has 33 significant bits.
    const double kInvPio2 = 0x1.45f306dc9c883p-1;
    const double kP1 = 0x1.921fb54400000p+0;
    const double kP1t = 0x1.0b4611a626331p-34;
    double k = 0.0;
    double r = x;
    if (pt_fabs_d(x) > 0x1.921fb54442d18p-1) {
        k = pt_rint_small(x * kInvPio2);
        r = (x - k * kP1) - k * kP1t;
    }
    const int64_t ki = (int64_t)k;
    const int q = (int)(ki & 3);
    const double sr = pt_kernel_sin(r);
    const double cr = pt_kernel_cos(r);
    switch (q) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

PT_HD double pt_sin(double x) { double s, c; pt_sincos(x, &s, &c); return s; }
PT_HD double pt_cos(double x) { double s, c; pt_sincos(x, &s, &c); return c; }

// asin(s) for 0 <= |s| <= 0.5: s + s*z*A(z), A from (2n)!/(4^n (n!)^2 (2n+1)), n = 1..26.
PT_HD double pt_kernel_asin(double s)
{
    const double z = s * s;
    double a = 0x1.1052bc5fa960ap-9;
    a = a * z + 0x1.208d3570ae5a6p-9;
    a = a * z + 0x1.3275586c5f2f0p-9;
    a = a * z + 0x1.464c0950f7d47p-9;
    a = a * z + 0x1.5c5f56efaaaabp-9;
    a = a * z + 0x1.750de64d7d05fp-9;
    a = a * z + 0x1.90cb77f60c7cep-9;
    a = a * z + 0x1.b026f57b13b14p-9;
    a = a * z + 0x1.d3d2a8e0dd67dp-9;
    a = a * z + 0x1.fcaf8fb6db6dbp-9;
    a = a * z + 0x1.15ee9d45d1746p-8;
    a = a * z + 0x1.31683bdef7bdfp-8;
    a = a * z + 0x1.51ba308d3dcb1p-8;
    a = a * z + 0x1.782dda12f684cp-8;
    a = a * z + 0x1.a6863d70a3d71p-8;
    a = a * z + 0x1.df3bd37a6f4dfp-8;
    a = a * z + 0x1.12ef3cf3cf3cfp-7;
    a = a * z + 0x1.3fde50d79435ep-7;
    a = a * z + 0x1.7a87878787878p-7;
    a = a * z + 0x1.c99999999999ap-7;
    a = a * z + 0x1.1c4ec4ec4ec4fp-6;
    a = a * z + 0x1.6e8ba2e8ba2e9p-6;
    a = a * z + 0x1.f1c71c71c71c7p-6;
    a = a * z + 0x1.6db6db6db6db7p-5;
    a = a * z + 0x1.3333333333333p-4;
    a = a * z + 0x1.5555555555555p-3;
    return s + (s * z) * a;
}

// Correctly rounded double square root (IEEE on both the host and gfx950).
PT_HD double pt_sqrt_d(double x) { return __builtin_sqrt(x); }

// acos(x) on [-1, 1]; NaN outside (as C99 acos).
PT_HD double pt_acos(double x)
{
    const double kPio2Hi = 0x1.921fb54442d18p+0;
    const double kPio2Lo = 0x1.1a62633145c07p-54;
    const double kPiHi = 0x1.921fb54442d18p+1;
    const double kPiLo = 0x1.1a62633145c07p-53;
    if (!(x == x)) return x;
    if (x > 1.0 || x < -1.0) return (x - x) / (x - x);
    if (x <= 0.5 && x >= -0.5) {
        return kPio2Hi - (pt_kernel_asin(x) - kPio2Lo);
    }
    if (x > 0.5) {
        const double z = (1.0 - x) * 0.5;
        return 2.0 * pt_kernel_asin(pt_sqrt_d(z));
    }
    const double z = (1.0 + x) * 0.5;
    return kPiHi - (2.0 * pt_kernel_asin(pt_sqrt_d(z)) - kPiLo);
}

// atan(u) for |u| <= tan(pi/8): u + u*z*T(z), T from (-1)^n/(2n+1), n = 1..24.
PT_HD double pt_kernel_atan(double u)
{
    const double z = u * u;
    double t = 0x1.4e5e0a72f0539p-6;
    t = t * z + -0x1.5c9882b931057p-6;
    t = t * z + 0x1.6c16c16c16c17p-6;
    t = t * z + -0x1.7d05f417d05f4p-6;
    t = t * z + 0x1.8f9c18f9c18fap-6;
    t = t * z + -0x1.a41a41a41a41ap-6;
    t = t * z + 0x1.bacf914c1bad0p-6;
    t = t * z + -0x1.d41d41d41d41dp-6;
    t = t * z + 0x1.f07c1f07c1f08p-6;
    t = t * z + -0x1.0842108421084p-5;
    t = t * z + 0x1.1a7b9611a7b96p-5;
    t = t * z + -0x1.2f684bda12f68p-5;
    t = t * z + 0x1.47ae147ae147bp-5;
    t = t * z + -0x1.642c8590b2164p-5;
    t = t * z + 0x1.8618618618618p-5;
    t = t * z + -0x1.af286bca1af28p-5;
    t = t * z + 0x1.e1e1e1e1e1e1ep-5;
    t = t * z + -0x1.1111111111111p-4;
    t = t * z + 0x1.3b13b13b13b14p-4;
    t = t * z + -0x1.745d1745d1746p-4;
    t = t * z + 0x1.c71c71c71c71cp-4;
    t = t * z + -0x1.2492492492492p-3;
    t = t * z + 0x1.999999999999ap-3;
    t = t * z + -0x1.5555555555555p-2;
    return u + (u * z) * t;
}

// atan(t) for 0 <= t <= 1.
PT_HD double pt_atan_unit(double t)
{
    const double kTanPio8 = 0x1.a827999fcef32p-2; // tan(pi/8) rounded
    const double kPio4Hi = 0x1.921fb54442d18p-1;
    const double kPio4Lo = 0x1.1a62633145c07p-55;
    if (t <= kTanPio8) return pt_kernel_atan(t);
    const double u = (t - 1.0) / (t + 1.0);
    return kPio4Hi + (pt_kernel_atan(u) + kPio4Lo);
}

PT_HD int pt_signbit_d(double x)
{
    union { double d; uint64_t u; } v;
    v.d = x;
    return (int)(v.u >> 63);
}

PT_HD double pt_copysign_d(double mag, double sgn)
{
    const double a = mag < 0.0 ? -mag : mag;
    return pt_signbit_d(sgn) ? -a : a;
}

// atan2(y, x) with the C99 special cases for zeros and infinities.
PT_HD double pt_atan2(double y, double x)
{
    const double kPio2Hi = 0x1.921fb54442d18p+0;
    const double kPio2Lo = 0x1.1a62633145c07p-54;
    const double kPiHi = 0x1.921fb54442d18p+1;
    const double kPiLo = 0x1.1a62633145c07p-53;
    const double kInf = __builtin_inf();
    if (!(x == x) || !(y == y)) return x + y;
    if (y == 0.0) {
        if (pt_signbit_d(x)) return pt_copysign_d(kPiHi, y);
        return y;
    }
    if (x == 0.0) return pt_copysign_d(kPio2Hi, y);
    const double ax = x < 0.0 ? -x : x;
    const double ay = y < 0.0 ? -y : y;
    if (ax == kInf || ay == kInf) {
        double a;
        if (ax == kInf && ay == kInf) {
            a = x < 0.0 ? 0x1.2d97c7f3321d2p+1 : 0x1.921fb54442d18p-1; // 3pi/4 : pi/4
        } else if (ax == kInf) {
            a = x < 0.0 ? kPiHi : 0.0;
        } else {
            a = kPio2Hi;
        }
        return pt_copysign_d(a, y);
    }
    double a;
    if (ay <= ax) {
        a = pt_atan_unit(ay / ax);
    } else {
        a = kPio2Hi - (pt_atan_unit(ax / ay) - kPio2Lo);
    }
    if (x < 0.0) a = kPiHi - (a - kPiLo);
    return y < 0.0 ? -a : a;
}
