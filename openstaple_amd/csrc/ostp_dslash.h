// Per-site algebra of the even-odd staggered Dirac operator, shared by the
// single-lattice kernels (dslash.hip) and the batched solver (batch.hip).
//
// D_eo / D_oe of OpenAcc/fermion_matrix.c:47-157 with the per-link algebra of
// OpenAcc/matvecmul.h:88-172.
#pragma once

#include "ostp_blas.h"

namespace ostp {

template <typename R> struct Cplx;
template <> struct Cplx<double> { using T = double2; };
template <> struct Cplx<float> { using T = float2; };

enum DslashMode : int { kPlain = 0, kMdagM = 1, kMdagMDot = 2 };


// FP32 phases are evaluated in double (eight per site; the kernel is
// bandwidth-bound) and rounded once.
__device__ __forceinline__ void sincos_t(double a, double* s, double* c) { phase_sincos(a, s, c); }
__device__ __forceinline__ void sincos_t(float a, float* s, float* c) {
    double ds, dc;
    phase_sincos((double)a, &ds, &dc);
    *s = (float)ds;
    *c = (float)dc;
}

// Links and phases are read once per launch.  Measured per 32^4 Dslash on
// MI355X: FP64 links with non-temporal loads and phases with plain loads
// 82.6 us (85.7 us with both non-temporal, 92.5 us with neither); FP32 with
// plain loads 36.3 us (50.0 us with non-temporal 64-bit loads).
__device__ __forceinline__ double2 ld_link(const double2* p) { return ld_nt(p); }
__device__ __forceinline__ double ld_link(const double* p) { return *p; }
__device__ __forceinline__ float2 ld_link(const float2* p) { return *p; }
__device__ __forceinline__ float ld_link(const float* p) { return *p; }

template <typename C> __device__ __forceinline__ C mk(double a, double b);
template <> __device__ __forceinline__ double2 mk<double2>(double a, double b) { return make_double2(a, b); }
template <> __device__ __forceinline__ float2 mk<float2>(double a, double b) { return make_float2((float)a, (float)b); }

// One hop: o += U(iu) e^{+i th} in(iv)  (DAG=false, mat_vec_mul_arg)
//          o -= U(iu)^dag e^{-i th} in(iv)  (DAG=true, conjmat_vec_mul_arg)
template <typename R, bool DAG>
__device__ __forceinline__ void hop(const typename Cplx<R>::T* __restrict__ U, const R* __restrict__ P,
                                    const typename Cplx<R>::T* __restrict__ in, int sizeh, int iu, int iv,
                                    typename Cplx<R>::T& o0, typename Cplx<R>::T& o1,
                                    typename Cplx<R>::T& o2) {
    using C = typename Cplx<R>::T;
    const C a0 = ld_link(U + iu), a1 = ld_link(U + sizeh + iu), a2 = ld_link(U + 2 * sizeh + iu);
    const C b0 = ld_link(U + 3 * sizeh + iu), b1 = ld_link(U + 4 * sizeh + iu), b2 = ld_link(U + 5 * sizeh + iu);
    const R arg = ld_link(P + iu);
    C v0 = in[iv], v1 = in[sizeh + iv], v2 = in[2 * sizeh + iv];
    R s, c;
    sincos_t(arg, &s, &c);
    const C phase = DAG ? mk<C>(c, -s) : mk<C>(c, s);
    v0 = cmul(v0, phase);
    v1 = cmul(v1, phase);
    v2 = cmul(v2, phase);
    // x = row0 x row1 ; row2 = conj(x)
    const C x0 = csub(cmul(a1, b2), cmul(a2, b1));
    const C x1 = csub(cmul(a2, b0), cmul(a0, b2));
    const C x2 = csub(cmul(a0, b1), cmul(a1, b0));
    if (!DAG) {
        C r0 = cmul(a0, v0); r0 = cfma(a1, v1, r0); r0 = cfma(a2, v2, r0);
        C r1 = cmul(b0, v0); r1 = cfma(b1, v1, r1); r1 = cfma(b2, v2, r1);
        C r2 = cmul(cconj(x0), v0); r2 = cfmac(x1, v1, r2); r2 = cfmac(x2, v2, r2);
        o0 = cadd(o0, r0); o1 = cadd(o1, r1); o2 = cadd(o2, r2);
    } else {
        C r0 = cmul(cconj(a0), v0); r0 = cfmac(b0, v1, r0); r0 = cfma(x0, v2, r0);
        C r1 = cmul(cconj(a1), v0); r1 = cfmac(b1, v1, r1); r1 = cfma(x1, v2, r1);
        C r2 = cmul(cconj(a2), v0); r2 = cfmac(b2, v1, r2); r2 = cfma(x2, v2, r2);
        o0 = csub(o0, r0); o1 = csub(o1, r1); o2 = csub(o2, r2);
    }
}

template <typename C> __device__ __forceinline__ double re_dot(C a, C b) {
    return (double)a.x * (double)b.x + (double)a.y * (double)b.y;
}

// Lattice extents as the kernels see them.
struct SiteGeom {
    int nd0h, nd1, nd2, nd3, nd0;
    int sizeh;
};

// out(idx) = 1/2 [ sum_mu U_mu(x) e^{i th} in(x+mu) - sum_mu U_mu(x-mu)^dag e^{-i th} in(x-mu) ]
// for the output half-site idx of parity EO (0: Deo, even output; 1: Doe).
// u[2*mu+parity], ph[2*mu+parity]: link planes and phases of the 8 arrays.
template <typename R, int EO>
__device__ __forceinline__ void dslash_site(const typename Cplx<R>::T* const* u, const R* const* ph,
                                            const typename Cplx<R>::T* __restrict__ in, const SiteGeom& a,
                                            int idx, typename Cplx<R>::T& o0, typename Cplx<R>::T& o1,
                                            typename Cplx<R>::T& o2) {
    using C = typename Cplx<R>::T;
    const int nd0h = a.nd0h, nd1 = a.nd1, nd2 = a.nd2, nd3 = a.nd3;
    int t = idx / nd0h;
    const int hd0 = idx - t * nd0h;
    int t2 = t / nd1;
    const int d1 = t - t2 * nd1;
    const int d3 = t2 / nd2;
    const int d2 = t2 - d3 * nd2;
    const int d0 = 2 * hd0 + ((d1 + d2 + d3 + EO) & 1);
    const int row = idx - hd0;
    const int s1 = nd0h, s2 = nd0h * nd1, s3 = s2 * nd2;
    const int ip0 = row + ((d0 + 1 == a.nd0) ? 0 : ((d0 + 1) >> 1));
    const int im0 = row + ((d0 == 0) ? ((a.nd0 - 1) >> 1) : ((d0 - 1) >> 1));
    const int ip1 = idx + (d1 == nd1 - 1 ? -(nd1 - 1) * s1 : s1);
    const int im1 = idx + (d1 == 0 ? (nd1 - 1) * s1 : -s1);
    const int ip2 = idx + (d2 == nd2 - 1 ? -(nd2 - 1) * s2 : s2);
    const int im2 = idx + (d2 == 0 ? (nd2 - 1) * s2 : -s2);
    const int ip3 = idx + (d3 == nd3 - 1 ? -(nd3 - 1) * s3 : s3);
    const int im3 = idx + (d3 == 0 ? (nd3 - 1) * s3 : -s3);
    constexpr int F = (EO == 0) ? 0 : 1;  // parity of forward link arrays
    constexpr int B = 1 - F;              // parity of backward link arrays
    o0 = mk<C>(0, 0);
    o1 = mk<C>(0, 0);
    o2 = mk<C>(0, 0);
    hop<R, true>(u[0 + B], ph[0 + B], in, a.sizeh, im0, im0, o0, o1, o2);
    hop<R, true>(u[2 + B], ph[2 + B], in, a.sizeh, im1, im1, o0, o1, o2);
    hop<R, true>(u[4 + B], ph[4 + B], in, a.sizeh, im2, im2, o0, o1, o2);
    hop<R, true>(u[6 + B], ph[6 + B], in, a.sizeh, im3, im3, o0, o1, o2);
    hop<R, false>(u[0 + F], ph[0 + F], in, a.sizeh, idx, ip0, o0, o1, o2);
    hop<R, false>(u[2 + F], ph[2 + F], in, a.sizeh, idx, ip1, o0, o1, o2);
    hop<R, false>(u[4 + F], ph[4 + F], in, a.sizeh, idx, ip2, o0, o1, o2);
    hop<R, false>(u[6 + F], ph[6 + F], in, a.sizeh, idx, ip3, o0, o1, o2);
    const R half = (R)0.5;
    o0.x *= half; o0.y *= half; o1.x *= half; o1.y *= half; o2.x *= half; o2.y *= half;
}

// Hop of several right-hand sides on one gauge field: the link (two rows,
// the rebuilt third, its phase) is loaded and prepared once and applied to
// each active system's spinor with the arithmetic of hop<R, DAG>.
template <typename R, bool DAG, int NR>
__device__ __forceinline__ void hop_mr(const typename Cplx<R>::T* __restrict__ U, const R* __restrict__ P,
                                       const typename Cplx<R>::T* const (&in)[NR], const bool (&act)[NR], int sizeh,
                                       int iu, int iv, typename Cplx<R>::T (&o)[NR][3]) {
    using C = typename Cplx<R>::T;
    const C a0 = ld_link(U + iu), a1 = ld_link(U + sizeh + iu), a2 = ld_link(U + 2 * sizeh + iu);
    const C b0 = ld_link(U + 3 * sizeh + iu), b1 = ld_link(U + 4 * sizeh + iu), b2 = ld_link(U + 5 * sizeh + iu);
    const R arg = ld_link(P + iu);
    R s, c;
    sincos_t(arg, &s, &c);
    const C phase = DAG ? mk<C>(c, -s) : mk<C>(c, s);
    const C x0 = csub(cmul(a1, b2), cmul(a2, b1));
    const C x1 = csub(cmul(a2, b0), cmul(a0, b2));
    const C x2 = csub(cmul(a0, b1), cmul(a1, b0));
#pragma unroll
    for (int r = 0; r < NR; r++) {
        if (!act[r]) continue;
        const C* v = in[r];
        const C v0 = cmul(v[iv], phase), v1 = cmul(v[sizeh + iv], phase), v2 = cmul(v[2 * sizeh + iv], phase);
        if (!DAG) {
            C r0 = cmul(a0, v0); r0 = cfma(a1, v1, r0); r0 = cfma(a2, v2, r0);
            C r1 = cmul(b0, v0); r1 = cfma(b1, v1, r1); r1 = cfma(b2, v2, r1);
            C r2 = cmul(cconj(x0), v0); r2 = cfmac(x1, v1, r2); r2 = cfmac(x2, v2, r2);
            o[r][0] = cadd(o[r][0], r0); o[r][1] = cadd(o[r][1], r1); o[r][2] = cadd(o[r][2], r2);
        } else {
            C r0 = cmul(cconj(a0), v0); r0 = cfmac(b0, v1, r0); r0 = cfma(x0, v2, r0);
            C r1 = cmul(cconj(a1), v0); r1 = cfmac(b1, v1, r1); r1 = cfma(x1, v2, r1);
            C r2 = cmul(cconj(a2), v0); r2 = cfmac(b2, v1, r2); r2 = cfma(x2, v2, r2);
            o[r][0] = csub(o[r][0], r0); o[r][1] = csub(o[r][1], r1); o[r][2] = csub(o[r][2], r2);
        }
    }
}

// Neighbour half-site indices of output half-site idx of parity EO (0 even,
// 1 odd): ip[mu] = idx + mu, im[mu] = idx - mu (periodic), as in dslash_site.
struct Nbrs {
    int ip[4], im[4];
};
template <int EO>
__device__ __forceinline__ Nbrs site_neighbours(const SiteGeom& a, int idx) {
    const int nd0h = a.nd0h, nd1 = a.nd1, nd2 = a.nd2, nd3 = a.nd3;
    int t = idx / nd0h;
    const int hd0 = idx - t * nd0h;
    int t2 = t / nd1;
    const int d1 = t - t2 * nd1;
    const int d3 = t2 / nd2;
    const int d2 = t2 - d3 * nd2;
    const int d0 = 2 * hd0 + ((d1 + d2 + d3 + EO) & 1);
    const int row = idx - hd0;
    const int s1 = nd0h, s2 = nd0h * nd1, s3 = s2 * nd2;
    Nbrs n;
    n.ip[0] = row + ((d0 + 1 == a.nd0) ? 0 : ((d0 + 1) >> 1));
    n.im[0] = row + ((d0 == 0) ? ((a.nd0 - 1) >> 1) : ((d0 - 1) >> 1));
    n.ip[1] = idx + (d1 == nd1 - 1 ? -(nd1 - 1) * s1 : s1);
    n.im[1] = idx + (d1 == 0 ? (nd1 - 1) * s1 : -s1);
    n.ip[2] = idx + (d2 == nd2 - 1 ? -(nd2 - 1) * s2 : s2);
    n.im[2] = idx + (d2 == 0 ? (nd2 - 1) * s2 : -s2);
    n.ip[3] = idx + (d3 == nd3 - 1 ? -(nd3 - 1) * s3 : s3);
    n.im[3] = idx + (d3 == 0 ? (nd3 - 1) * s3 : -s3);
    return n;
}

}  // namespace ostp
