// acos for the one window-dependent angle of the pseudo-source sweep
// (ICHDevice.cu:815-819).  Method: fdlibm e_acos.c (Sun Microsystems, 1993):
// a rational approximation R(z) of (asin(x) - x) / x^3 on |x| <= 0.5 and the
// half-angle reductions acos(x) = pi - 2 asin(sqrt((1+x)/2)) for x < -0.5 and
// acos(x) = 2 asin(sqrt((1-x)/2)) for x > 0.5.  Written with plain IEEE
// operations only (+ - * / sqrt), so the host and the device produce the same
// bits when both are compiled with -ffp-contract=off.  Error < 1 ulp against
// glibc's acos.
//
// The coefficients and reductions follow fdlibm's e_acos.c, whose notice is
// preserved as its licence asks:
//   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//   Developed at SunPro, a Sun Microsystems, Inc. business.
//   Permission to use, copy, modify, and distribute this software is freely
//   granted, provided that this notice is preserved.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>

namespace svg {

__host__ __device__ __forceinline__ uint64_t dbits(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint64_t)__double_as_longlong(x);
#else
    uint64_t b; std::memcpy(&b, &x, 8); return b;
#endif
}

__host__ __device__ __forceinline__ double bitsd(uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __longlong_as_double((long long)b);
#else
    double x; std::memcpy(&x, &b, 8); return x;
#endif
}

__host__ __device__ __forceinline__ double acos_ratR(double z)
{
    const double pS0 = 1.66666666666666657415e-01, pS1 = -3.25565818622400915405e-01,
                 pS2 = 2.01212532134862925881e-01, pS3 = -4.00555345006794114027e-02,
                 pS4 = 7.91534994289814532176e-04, pS5 = 3.47933107596021167570e-05,
                 qS1 = -2.40339491173441421878e+00, qS2 = 2.02094576023350569471e+00,
                 qS3 = -6.88283971605453293030e-01, qS4 = 7.70381505559019352791e-02;
    double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    double q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    return p / q;
}

__host__ __device__ inline double svg_acos(double x)
{
    const double pi = 3.14159265358979311600e+00;
    const double pio2_hi = 1.57079632679489655800e+00;
    const double pio2_lo = 6.12323399573676603587e-17;
    const uint64_t bits = dbits(x);
    const uint32_t ix = (uint32_t)(bits >> 32) & 0x7fffffffu;
    if (ix >= 0x3ff00000u) {
        const uint32_t lo = (uint32_t)bits;
        if (((ix - 0x3ff00000u) | lo) == 0) {
            if ((int64_t)bits > 0) return 0.0;
            return pi + 2.0 * pio2_lo;
        }
        return (x - x) / (x - x);
    }
    if (ix < 0x3fe00000u) {
        if (ix <= 0x3c600000u) return pio2_hi + pio2_lo;
        const double z = x * x;
        const double r = acos_ratR(z);
        return pio2_hi - (x - (pio2_lo - x * r));
    }
    if ((int64_t)bits < 0) {
        const double z = (1.0 + x) * 0.5;
        const double s = sqrt(z);
        const double r = acos_ratR(z);
        const double w = r * s - pio2_lo;
        return pi - 2.0 * (s + w);
    }
    const double z = (1.0 - x) * 0.5;
    const double s = sqrt(z);
    const double df = bitsd(dbits(s) & 0xFFFFFFFF00000000ull);
    const double c = (z - df * df) / (s + df);
    const double r = acos_ratR(z);
    const double w = r * s + c;
    return 2.0 * (df + w);
}

}  // namespace svg
