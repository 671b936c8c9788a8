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
#include "ostp_dslash.h"

namespace ostp {

template <typename R>
struct DslashArgs {
    using C = typename Cplx<R>::T;
    const C* u[8];      // start of r0.c0 of link array 2*mu+parity
    const R* ph[8];     // phase array 2*mu+parity
    const C* in;
    C* out;
    const C* x;         // epilogue input: out = m2*x - D in
    double m2;
    SiteGeom g;
    int tmajor;         // workgroup order: t-slices fastest (see k_dslash)
    int rev;            // each XCD walks its slab backwards
    SolverCtl ctl;
};

// EO = 0: Deo (even output, odd input); EO = 1: Doe.
template <typename R, int EO, int MODE>
__global__ __launch_bounds__(kBlock) void k_dslash(DslashArgs<R> a) {
    using C = typename Cplx<R>::T;
    if (!pred_pass(a.ctl)) return;
    // Workgroup -> sites: the XCD remap gives each XCD a contiguous range of
    // logical tiles; with tmajor the logical tile index runs over t fastest,
    // so the +-t neighbour slices of a tile are streamed by the same XCD at
    // nearly the same time and its L2 serves their re-reads (FP64: 81.2 us vs
    // 85.5 us per 32^4 Dslash).
    const int blk = xcd_block_id_dir(a.rev);
    int idx;
    if (a.tmajor) {
        const int s3 = a.g.sizeh / a.g.nd3;
        const int tile = blk / a.g.nd3, tt = blk - tile * a.g.nd3;
        idx = tt * s3 + tile * kBlock + (int)threadIdx.x;
    } else {
        idx = blk * kBlock + (int)threadIdx.x;
    }
    double dot = 0.0;
    if (idx < a.g.sizeh) {
        C o0, o1, o2;
        dslash_site<R, EO>(a.u, a.ph, a.in, a.g, idx, o0, o1, o2);
        if (MODE == kPlain) {
            if constexpr (sizeof(R) == 8) {
                st_nt(a.out + idx, o0);
                st_nt(a.out + a.g.sizeh + idx, o1);
                st_nt(a.out + 2 * a.g.sizeh + idx, o2);
            } else {
                a.out[idx] = o0;
                a.out[a.g.sizeh + idx] = o1;
                a.out[2 * a.g.sizeh + idx] = o2;
            }
        } else {
            // combine_in1xferm_mass_minus_in2: out = x*m2 - D in  (promoted to double)
            const C x0 = a.x[idx], x1 = a.x[a.g.sizeh + idx], x2 = a.x[2 * a.g.sizeh + idx];
            const double m2 = a.m2;
            const C s0 = mk<C>((double)x0.x * m2 - (double)o0.x, (double)x0.y * m2 - (double)o0.y);
            const C s1v = mk<C>((double)x1.x * m2 - (double)o1.x, (double)x1.y * m2 - (double)o1.y);
            const C s2v = mk<C>((double)x2.x * m2 - (double)o2.x, (double)x2.y * m2 - (double)o2.y);
            if constexpr (sizeof(R) == 8) {
                st_nt(a.out + idx, s0);
                st_nt(a.out + a.g.sizeh + idx, s1v);
                st_nt(a.out + 2 * a.g.sizeh + idx, s2v);
            } else {
                a.out[idx] = s0;
                a.out[a.g.sizeh + idx] = s1v;
                a.out[2 * a.g.sizeh + idx] = s2v;
            }
            if (MODE == kMdagMDot) dot = re_dot(x0, s0) + re_dot(x1, s1v) + re_dot(x2, s2v);
        }
    }
    // partial slot = logical tile, so the reduction order is the same in both walk directions
    if (MODE == kMdagMDot) block_partial_at<kBlock>(dot, 0.0, a.ctl.partials, blk);
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
    a.g.nd0 = L.nd[0];
    a.g.nd0h = L.nd[0] / 2;
    a.g.nd1 = L.nd[1];
    a.g.nd2 = L.nd[2];
    a.g.nd3 = L.nd[3];
    a.g.sizeh = (int)L.sizeh;
    // t-fastest tiles need whole tiles per t-slice; used for FP64 only (no gain in FP32)
    const long s3 = L.sizeh / L.nd[3];
    a.tmajor = (sizeof(R) == 8 && s3 % kBlock == 0) ? 1 : 0;
    a.rev = 0;
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
    a.rev = ctx().dslash_serpentine ? (int)(ctx().dslash_launches++ & 1u) : 0;
    const int nb = blocks_for(ctx().lat.sizeh);
    hipStream_t s = ctx().stream;
#define OSTP_L(EO, MODE) hipLaunchKernelGGL((k_dslash<R, EO, MODE>), dim3(nb), dim3(kBlock), 0, s, a)
    if (mode == kPlain) {
        if (eo == 0) OSTP_L(0, kPlain); else OSTP_L(1, kPlain);
    } else if (mode == kMdagM) {
        if (eo == 0) OSTP_L(0, kMdagM); else OSTP_L(1, kMdagM);
    } else {
        if (eo != 0) return fail("fused dot epilogue only on Deo");
        if (fin != kFinStore && fin != kFinCgAlpha && fin != kFinMsAlpha && fin != kFinMxAlpha &&
            fin != kFinFcgAlpha && fin != kFinMxmAlpha)
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
static bool sync_ok(const char* what) { return entry_done(what); }

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
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double us = -1.0;
    if (hip_ok(hipEventCreate(&e0), "event") && hip_ok(hipEventCreate(&e1), "event") &&
        hip_ok(hipEventRecord(e0, ctx().stream), "event record")) {
        bool ok = true;
        for (int r = 0; r < reps && ok; r++)
            ok = single ? launch_dslash<float>(eo, kPlain, kFinStore, du, dp, di, dout, nullptr, 0.0, SolverCtl{})
                        : launch_dslash<double>(eo, kPlain, kFinStore, du, dp, di, dout, nullptr, 0.0, SolverCtl{});
        float ms = 0.f;
        if (ok && hip_ok(hipEventRecord(e1, ctx().stream), "event record") &&
            hip_ok(hipEventSynchronize(e1), "event sync") && hip_ok(hipEventElapsedTime(&ms, e0, e1), "elapsed"))
            us = 1000.0 * (double)ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return us;
}

}  // extern "C"
