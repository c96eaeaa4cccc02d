// flr_sincos.hpp -- sine and cosine of a galactic-binary phase (|x| up to ~1e8 rad).
//
// Phases of year-long sources run to millions of radians. The argument is reduced by pi/2
// (Cody and Waite's method: n = nearest integer to x * 2/pi, then r = x - n*pi/2 with pi/2 split
// into three doubles P1 + P2 + P3, each product subtracted with one fused multiply-add), and sin
// and cos of r in [-pi/4, pi/4] come from their Taylor series (through r^17 and r^16: the first
// omitted terms are below 1e-19). For |x| < 2^28 the reduced argument is within ~1e-16 of the
// exact x - n*pi/2 and the results within ~2e-16 of sin x, cos x; larger or non-finite
// arguments take the math library's sincos. Written for host and device: the host build (the
// accuracy test) defines FLR_HOST_DEVICE empty.
#pragma once

#include <cmath>

#ifndef FLR_HOST_DEVICE
#define FLR_HOST_DEVICE __host__ __device__
#endif

namespace flr {

FLR_HOST_DEVICE inline void phase_sincos(double x, double *s, double *c)
{
    if (!(fabs(x) < 268435456.0)) {  // 2^28 (and NaN)
        sincos(x, s, c);
        return;
    }
    const double two_over_pi = 0x1.45f306dc9c883p-1;
    const double P1 = 0x1.921fb54442d18p+0;    // pi/2 rounded to double
    const double P2 = 0x1.1a62633145c07p-54;   // pi/2 - P1, rounded
    const double P3 = -0x1.f1976b7ed8fbcp-110; // pi/2 - P1 - P2, rounded
    const double n = rint(x * two_over_pi);
    double r = fma(-n, P1, x);
    r = fma(-n, P2, r);
    r = fma(-n, P3, r);
    const double z = r * r;
    double ps = 1.0 / 355687428096000.0;  // 1/17!
    ps = fma(ps, z, -1.0 / 1307674368000.0);
    ps = fma(ps, z, 1.0 / 6227020800.0);
    ps = fma(ps, z, -1.0 / 39916800.0);
    ps = fma(ps, z, 1.0 / 362880.0);
    ps = fma(ps, z, -1.0 / 5040.0);
    ps = fma(ps, z, 1.0 / 120.0);
    ps = fma(ps, z, -1.0 / 6.0);
    const double sr = fma(ps * z, r, r);
    double pc = 1.0 / 20922789888000.0;  // 1/16!
    pc = fma(pc, z, -1.0 / 87178291200.0);
    pc = fma(pc, z, 1.0 / 479001600.0);
    pc = fma(pc, z, -1.0 / 3628800.0);
    pc = fma(pc, z, 1.0 / 40320.0);
    pc = fma(pc, z, -1.0 / 720.0);
    pc = fma(pc, z, 1.0 / 24.0);
    pc = fma(pc, z, -0.5);
    const double cr = fma(pc, z, 1.0);
    // x = n pi/2 + r: quadrant q = n mod 4 rotates (cos r, sin r) by q quarter turns
    const int q = (int)((long long)n & 3);
    const double sq = (q & 1) ? cr : sr;
    const double cq = (q & 1) ? sr : cr;
    *s = (q & 2) ? -sq : sq;
    *c = ((q + 1) & 2) ? -cq : cq;
}

}  // namespace flr
