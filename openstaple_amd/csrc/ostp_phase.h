// Phase factor e^{i theta} = (cos theta, sin theta) of the U(1) links.
// Plain C/C++ so that host tests can compile it with gcc; HIP compiles it for
// host and device.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#define OSTP_PHASE_FN __host__ __device__ __forceinline__
#else
#define OSTP_PHASE_FN static inline
#endif

// sin and cos of a U(1) phase angle.
//
// Argument reduction x = n*pi/2 + r is Cody-Waite with pi/2 split in three
// doubles and fused multiply-adds; sin r and cos r on |r| <= pi/4 are the
// minimax polynomials of fdlibm's k_sin.c / k_cos.c (Sun Microsystems, 1993,
// freely redistributable).  Absolute error <= 1.2e-16 against the C library's
// sin/cos for |x| <= 1e8 (phases of a background field are far below that);
// at multiples of pi/2 (staggered signs, antiperiodic boundaries) the unit
// component is exact and the near-zero one within 1e-27.  The math library's
// sincos carries a Payne-Hanek path for huge arguments whose register demand
// cuts the occupancy of a kernel that calls it eight times per site from
// four waves per SIMD to one.
OSTP_PHASE_FN void phase_sincos(double x, double* s, double* c) {
    const double n = rint(x * 0.63661977236758134308);
    double r = fma(-n, 1.5707963267948966, x);
    r = fma(-n, 6.123233995736766e-17, r);
    r = fma(-n, -1.4973849048591698e-33, r);
    const double z = r * r, w = z * z;
    const double ps = 8.33333333332248946124e-03 + z * (-1.98412698298579493134e-04 + z * 2.75573137070700676789e-06) +
                      z * w * (-2.50507602534068634195e-08 + z * 1.58969099521155010221e-10);
    const double sn = r + z * r * (-1.66666666666666324348e-01 + z * ps);
    const double pc = z * (4.16666666666666019037e-02 + z * (-1.38888888888741095749e-03 + z * 2.48015872894767294178e-05)) +
                      w * w * (-2.75573143513906633035e-07 + z * (2.08757232129817482790e-09 + z * -1.13596475577881948265e-11));
    const double hz = 0.5 * z, one_m_hz = 1.0 - hz;
    const double cs = one_m_hz + (((1.0 - one_m_hz) - hz) + z * pc);
    const int q = ((int)(long long)n) & 3;
    double ss = (q & 1) ? cs : sn, cc = (q & 1) ? sn : cs;
    if (q == 1 || q == 2) cc = -cc;
    if (q >= 2) ss = -ss;
    *s = ss;
    *c = cc;
}
