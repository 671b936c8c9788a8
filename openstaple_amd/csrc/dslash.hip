// Even-odd staggered Dirac operator on gfx950.
//
// D_eo / D_oe of OpenAcc/fermion_matrix.c:47-157 with the per-link algebra of
// OpenAcc/matvecmul.h:88-172: for an output half-site x of one parity,
//   out(x) = 1/2 [ sum_mu U_mu(x) e^{i th_mu(x)} in(x+mu)
//                - sum_mu U_mu(x-mu)^dag e^{-i th_mu(x-mu)} in(x-mu) ],
// with only rows 0,1 of each link read (row 2 = conj(row0 x row1)) and the
// U(1) phase e^{i th} = cos th + i sin th folded in at load time.
//
// One thread per output half-site; consecutive lanes own consecutive
// half-sites, so every SoA plane access is a fully coalesced 128-bit (FP64)
// or 64-bit (FP32) load.  Workgroups are remapped so each XCD streams a
// contiguous slab of the lattice and its L2 serves the neighbour-spinor
// re-reads.  The kernel is HBM-bound (about 0.6 flop/byte); no MFMA.
//
// Epilogues fuse fermion_matrix_multiplication_shifted's
// combine_in1xferm_mass_minus_in2 (fermionic_utilities.c:261-272) and the
// CG's alpha = (p, s) reduction (inverter_full.c:76) into the second hop.
#include "ostp_blas.h"

namespace ostp {

template <typename R> struct Cplx;
template <> struct Cplx<double> { using T = double2; };
template <> struct Cplx<float> { using T = float2; };

enum DslashMode : int { kPlain = 0, kMdagM = 1, kMdagMDot = 2 };

template <typename R>
struct DslashArgs {
    using C = typename Cplx<R>::T;
    const C* u[8];      // start of r0.c0 of link array 2*mu+parity
    const R* ph[8];     // phase array 2*mu+parity
    const C* in;
    C* out;
    const C* x;         // epilogue input: out = m2*x - D in
    double m2;
    int nd0h, nd1, nd2, nd3, nd0;
    int sizeh;
    SolverCtl ctl;
};

// FP32 phases are evaluated in double (eight per site; the kernel is
// bandwidth-bound) and rounded once.
__device__ __forceinline__ void sincos_t(double a, double* s, double* c) { phase_sincos(a, s, c); }
__device__ __forceinline__ void sincos_t(float a, float* s, float* c) {
    double ds, dc;
    phase_sincos((double)a, &ds, &dc);
    *s = (float)ds;
    *c = (float)dc;
}

// Links and phases are read once per launch.  FP64: non-temporal loads
// (measured 85.7 us vs 92.5 us per 32^4 Dslash); FP32: plain loads (36.3 us
// vs 50.0 us with non-temporal 64-bit loads).
__device__ __forceinline__ double2 ld_link(const double2* p) { return ld_nt(p); }
__device__ __forceinline__ double ld_link(const double* p) { return ld_nt(p); }
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

// EO = 0: Deo (even output, odd input); EO = 1: Doe.
template <typename R, int EO, int MODE>
__global__ __launch_bounds__(kBlock) void k_dslash(DslashArgs<R> a) {
    using C = typename Cplx<R>::T;
    if (!pred_pass(a.ctl)) return;
    const int idx = xcd_block_id() * kBlock + (int)threadIdx.x;
    double dot = 0.0;
    if (idx < a.sizeh) {
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
        C o0 = mk<C>(0, 0), o1 = mk<C>(0, 0), o2 = mk<C>(0, 0);
        hop<R, true>(a.u[0 + B], a.ph[0 + B], a.in, a.sizeh, im0, im0, o0, o1, o2);
        hop<R, true>(a.u[2 + B], a.ph[2 + B], a.in, a.sizeh, im1, im1, o0, o1, o2);
        hop<R, true>(a.u[4 + B], a.ph[4 + B], a.in, a.sizeh, im2, im2, o0, o1, o2);
        hop<R, true>(a.u[6 + B], a.ph[6 + B], a.in, a.sizeh, im3, im3, o0, o1, o2);
        hop<R, false>(a.u[0 + F], a.ph[0 + F], a.in, a.sizeh, idx, ip0, o0, o1, o2);
        hop<R, false>(a.u[2 + F], a.ph[2 + F], a.in, a.sizeh, idx, ip1, o0, o1, o2);
        hop<R, false>(a.u[4 + F], a.ph[4 + F], a.in, a.sizeh, idx, ip2, o0, o1, o2);
        hop<R, false>(a.u[6 + F], a.ph[6 + F], a.in, a.sizeh, idx, ip3, o0, o1, o2);
        const R half = (R)0.5;
        o0.x *= half; o0.y *= half; o1.x *= half; o1.y *= half; o2.x *= half; o2.y *= half;
        if (MODE == kPlain) {
            a.out[idx] = o0;
            a.out[a.sizeh + idx] = o1;
            a.out[2 * a.sizeh + idx] = o2;
        } else {
            // combine_in1xferm_mass_minus_in2: out = x*m2 - D in  (promoted to double)
            const C x0 = a.x[idx], x1 = a.x[a.sizeh + idx], x2 = a.x[2 * a.sizeh + idx];
            const double m2 = a.m2;
            const C s0 = mk<C>((double)x0.x * m2 - (double)o0.x, (double)x0.y * m2 - (double)o0.y);
            const C s1v = mk<C>((double)x1.x * m2 - (double)o1.x, (double)x1.y * m2 - (double)o1.y);
            const C s2v = mk<C>((double)x2.x * m2 - (double)o2.x, (double)x2.y * m2 - (double)o2.y);
            a.out[idx] = s0;
            a.out[a.sizeh + idx] = s1v;
            a.out[2 * a.sizeh + idx] = s2v;
            if (MODE == kMdagMDot) dot = re_dot(x0, s0) + re_dot(x1, s1v) + re_dot(x2, s2v);
        }
    }
    if (MODE == kMdagMDot) block_partial<1>(dot, 0.0, a.ctl);
}

// ---------------------------------------------------------------- launchers

template <typename R>
static bool fill_args(DslashArgs<R>& a, const void* u_dev, const void* ph_dev, const void* in_dev, void* out_dev) {
    const Lattice& L = ctx().lat;
    using C = typename Cplx<R>::T;
    const size_t su3 = sizeof(R) == 8 ? L.su3_bytes() : L.su3f_bytes();
    const size_t soa = sizeof(R) == 8 ? L.dsoa_bytes() : L.fsoa_bytes();
    for (int l = 0; l < 8; l++) {
        a.u[l] = (const C*)((const char*)u_dev + l * su3);
        a.ph[l] = (const R*)((const char*)ph_dev + l * soa);
    }
    a.in = (const C*)in_dev;
    a.out = (C*)out_dev;
    a.x = nullptr;
    a.m2 = 0.0;
    a.nd0 = L.nd[0];
    a.nd0h = L.nd[0] / 2;
    a.nd1 = L.nd[1];
    a.nd2 = L.nd[2];
    a.nd3 = L.nd[3];
    a.sizeh = (int)L.sizeh;
    return true;
}

// Launch one Dslash on device pointers.  mode/fin select the fused epilogue.
template <typename R>
bool launch_dslash(int eo, int mode, int fin, const void* u, const void* ph, const void* in, void* out,
                   const void* x, double m2, const SolverCtl& ctl) {
    DslashArgs<R> a;
    fill_args<R>(a, u, ph, in, out);
    a.x = (const typename Cplx<R>::T*)x;
    a.m2 = m2;
    a.ctl = ctl;
    const int nb = blocks_for(ctx().lat.sizeh);
    hipStream_t s = ctx().stream;
#define OSTP_L(EO, MODE) hipLaunchKernelGGL((k_dslash<R, EO, MODE>), dim3(nb), dim3(kBlock), 0, s, a)
    if (mode == kPlain) {
        if (eo == 0) OSTP_L(0, kPlain); else OSTP_L(1, kPlain);
    } else if (mode == kMdagM) {
        if (eo == 0) OSTP_L(0, kMdagM); else OSTP_L(1, kMdagM);
    } else {
        if (eo != 0) return fail("fused dot epilogue only on Deo");
        if (fin != kFinStore && fin != kFinCgAlpha && fin != kFinMsAlpha && fin != kFinMxAlpha)
            return fail("unsupported dslash finalizer %d", fin);
        OSTP_L(0, kMdagMDot);
    }
#undef OSTP_L
    if (!hip_ok(hipGetLastError(), "k_dslash launch")) return false;
    return mode == kMdagMDot ? launch_finalize(fin, ctl, nb) : true;
}

template bool launch_dslash<double>(int, int, int, const void*, const void*, const void*, void*, const void*,
                                    double, const SolverCtl&);
template bool launch_dslash<float>(int, int, int, const void*, const void*, const void*, void*, const void*,
                                   double, const SolverCtl&);

// M = (m^2+shift) - Deo Doe on device pointers; temp receives Doe in.
template <typename R>
bool launch_mdagm(const void* u, const void* ph, const void* in, void* temp, void* out, double m2, int mode,
                  int fin, const SolverCtl& ctl) {
    if (!launch_dslash<R>(1, kPlain, kFinStore, u, ph, in, temp, nullptr, 0.0, ctl)) return false;
    return launch_dslash<R>(0, mode, fin, u, ph, temp, out, in, m2, ctl);
}

template bool launch_mdagm<double>(const void*, const void*, const void*, void*, void*, double, int, int,
                                   const SolverCtl&);
template bool launch_mdagm<float>(const void*, const void*, const void*, void*, void*, double, int, int,
                                  const SolverCtl&);

}  // namespace ostp

using namespace ostp;

// ------------------------------------------------------------ C entry points
static bool sync_ok(const char* what) { return hip_ok(hipStreamSynchronize(ctx().stream), what); }

extern "C" {

static void dslash_entry(int eo, const void* u, void* out, const void* in, const void* backfield, int single,
                         const char* name) {
    if (!require_init(name)) return;
    const Lattice& L = ctx().lat;
    const size_t vb = single ? L.vec3f_bytes() : L.vec3_bytes();
    const size_t ub = 8 * (single ? L.su3f_bytes() : L.su3_bytes());
    const size_t pb = 8 * (single ? L.fsoa_bytes() : L.dsoa_bytes());
    const void* du = resolve(u, ub, name);
    const void* dp = resolve(backfield, pb, name);
    const void* di = resolve(in, vb, name);
    void* dout = resolve(out, vb, name);
    if (!du || !dp || !di || !dout) return;
    bool ok = single ? launch_dslash<float>(eo, kPlain, kFinStore, du, dp, di, dout, nullptr, 0.0, SolverCtl{})
                     : launch_dslash<double>(eo, kPlain, kFinStore, du, dp, di, dout, nullptr, 0.0, SolverCtl{});
    if (ok) sync_ok(name);
}

void acc_Deo(const su3_soa* u, vec3_soa* out, const vec3_soa* in, const double_soa* backfield) {
    dslash_entry(0, u, out, in, backfield, 0, "acc_Deo");
}
void acc_Doe(const su3_soa* u, vec3_soa* out, const vec3_soa* in, const double_soa* backfield) {
    dslash_entry(1, u, out, in, backfield, 0, "acc_Doe");
}
void acc_Deo_f(const su3_soa_f* u, vec3_soa_f* out, const vec3_soa_f* in, const float_soa* backfield) {
    dslash_entry(0, u, out, in, backfield, 1, "acc_Deo_f");
}
void acc_Doe_f(const su3_soa_f* u, vec3_soa_f* out, const vec3_soa_f* in, const float_soa* backfield) {
    dslash_entry(1, u, out, in, backfield, 1, "acc_Doe_f");
}

static void mdagm_entry(const void* u, void* out, const void* in, void* temp, const ferm_param* pars,
                        double shift, int single, const char* name) {
    if (!require_init(name)) return;
    if (!pars) { fail("%s: null ferm_param", name); return; }
    const Lattice& L = ctx().lat;
    const size_t vb = single ? L.vec3f_bytes() : L.vec3_bytes();
    const size_t ub = 8 * (single ? L.su3f_bytes() : L.su3_bytes());
    const size_t pb = 8 * (single ? L.fsoa_bytes() : L.dsoa_bytes());
    const void* ph = single ? (const void*)pars->phases_f : (const void*)pars->phases;
    const void* du = resolve(u, ub, name);
    const void* dp = resolve(ph, pb, name);
    const void* di = resolve(in, vb, name);
    void* dt = resolve(temp, vb, name);
    void* dout = resolve(out, vb, name);
    if (!du || !dp || !di || !dt || !dout) return;
    const double m2 = single ? pars->ferm_mass * pars->ferm_mass + (double)(float)shift
                             : pars->ferm_mass * pars->ferm_mass + shift;
    bool ok = single ? launch_mdagm<float>(du, dp, di, dt, dout, m2, kMdagM, kFinStore, SolverCtl{})
                     : launch_mdagm<double>(du, dp, di, dt, dout, m2, kMdagM, kFinStore, SolverCtl{});
    if (ok) sync_ok(name);
}

void fermion_matrix_multiplication(const su3_soa* u, vec3_soa* out, const vec3_soa* in, vec3_soa* temp1,
                                   ferm_param* pars) {
    mdagm_entry(u, out, in, temp1, pars, 0.0, 0, "fermion_matrix_multiplication");
}
void fermion_matrix_multiplication_shifted(const su3_soa* u, vec3_soa* out, const vec3_soa* in, vec3_soa* temp1,
                                           ferm_param* pars, double shift) {
    mdagm_entry(u, out, in, temp1, pars, shift, 0, "fermion_matrix_multiplication_shifted");
}
void fermion_matrix_multiplication_shifted_f(const su3_soa_f* u, vec3_soa_f* out, const vec3_soa_f* in,
                                             vec3_soa_f* temp1, ferm_param* pars, double shift) {
    mdagm_entry(u, out, in, temp1, pars, shift, 1, "fermion_matrix_multiplication_shifted_f");
}

double ostp_time_dslash(const void* u, void* out, const void* in, const void* phases, int eo, int single,
                        int reps) {
    if (!require_init("ostp_time_dslash")) return -1.0;
    const Lattice& L = ctx().lat;
    const size_t vb = single ? L.vec3f_bytes() : L.vec3_bytes();
    const void* du = resolve(u, 8 * (single ? L.su3f_bytes() : L.su3_bytes()), "ostp_time_dslash");
    const void* dp = resolve(phases, 8 * (single ? L.fsoa_bytes() : L.dsoa_bytes()), "ostp_time_dslash");
    const void* di = resolve(in, vb, "ostp_time_dslash");
    void* dout = resolve(out, vb, "ostp_time_dslash");
    if (!du || !dp || !di || !dout || reps <= 0) return -1.0;
    hipEvent_t e0, e1;
    if (!hip_ok(hipEventCreate(&e0), "event") || !hip_ok(hipEventCreate(&e1), "event")) return -1.0;
    (void)hipEventRecord(e0, ctx().stream);
    for (int r = 0; r < reps; r++) {
        bool ok = single ? launch_dslash<float>(eo, kPlain, kFinStore, du, dp, di, dout, nullptr, 0.0, SolverCtl{})
                         : launch_dslash<double>(eo, kPlain, kFinStore, du, dp, di, dout, nullptr, 0.0, SolverCtl{});
        if (!ok) return -1.0;
    }
    (void)hipEventRecord(e1, ctx().stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 1000.0 * (double)ms / reps;
}

}  // extern "C"
