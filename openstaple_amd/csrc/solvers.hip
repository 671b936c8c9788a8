// Krylov solvers of the fermion hot path with device-resident scalars.
//
//  ker_invert_openacc        OpenAcc/inverter_full.c:19-132      (CG, restarts, SAFETY 0.95)
//  inverter_mixed_precision  OpenAcc/inverter_mixedp.c:41-181    (FP32 CG + FP64 reliable update, SAFETY 0.9)
//  multishift_invert         OpenAcc/inverter_multishift_full.c:23-252 (CG-M, per-shift flags, SAFETY 0.95)
//  recombine_shifted_vec3_to_vec3                                :254-282
//  inverter_wrapper / inverter_multishift_wrapper  OpenAcc/inverter_wrappers.c:45-159
//
// The reference returns every dot product to the host (two device->host syncs
// per iteration).  Here each reduction ends in an on-device finalizer that
// applies the same scalar recurrence, and every iteration's kernels carry a
// predicate (solver still running / this iteration started), so the host
// queues iterations in batches and synchronises only between batches.  The
// sequence of floating-point scalar operations is the reference's; vector
// updates are fused (x and r together; the multishift x_k and p_k together
// with p) without changing what each element computes.
#include "ostp_blas.h"
#include "ostp_dslash.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <chrono>
#include <functional>
#include <vector>

namespace ostp {

// ------------------------------------------------------------------ kernels
// x += omega p ; r -= omega s ; lambda = |r|^2   (inverter_full.c:82-86)
template <typename C>
__global__ __launch_bounds__(kBlock) void k_cg_xr(C* __restrict__ x, C* __restrict__ r, const C* __restrict__ p,
                                                  const C* __restrict__ s, int sizeh, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const double om = ctl.st->omega;
    double v = 0.0;
    if (i < sizeh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = (long)c * sizeh + i;
            const C pp = p[k], ss = s[k], xx = x[k], rr = r[k];
            x[k] = mkc<C>((double)pp.x * om + (double)xx.x, (double)pp.y * om + (double)xx.y);
            const C rn = mkc<C>((double)ss.x * (-om) + (double)rr.x, (double)ss.y * (-om) + (double)rr.y);
            r[k] = rn;
            v += (double)rn.x * rn.x + (double)rn.y * rn.y;
        }
    }
    block_partial<1>(v, 0.0, ctl);
}

// p = p*gammag + r   (inverter_full.c:96)
template <typename C>
__global__ __launch_bounds__(kBlock) void k_cg_p(C* __restrict__ p, const C* __restrict__ r, long n3, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n3) return;
    const double g = ctl.st->gammag;
    const C pp = p[i], rr = r[i];
    p[i] = mkc<C>((double)pp.x * g + (double)rr.x, (double)pp.y * g + (double)rr.y);
}

// r = b - s ; p = r ; |r|^2 -> finalizer F   (inverter_full.c:56-63)
template <int F>
__global__ __launch_bounds__(kBlock) void k_cg_restart(const double2* __restrict__ b, const double2* __restrict__ s,
                                                       double2* __restrict__ r, double2* __restrict__ p, int sizeh,
                                                       SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v = 0.0;
    if (i < sizeh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = (long)c * sizeh + i;
            const double2 bb = b[k], ss = s[k];
            const double2 rr = make_double2(bb.x - ss.x, bb.y - ss.y);
            r[k] = rr;
            if (p) p[k] = rr;
            v += rr.x * rr.x + rr.y * rr.y;
        }
    }
    block_partial<1>(v, 0.0, ctl);
}

// r += omega s ; lambda = |r|^2   (inverter_multishift_full.c:141-142)
template <typename C>
__global__ __launch_bounds__(kBlock) void k_ms_r(C* __restrict__ r, const C* __restrict__ s, int sizeh, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const double om = ctl.st->omega;
    double v = 0.0;
    if (i < sizeh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = (long)c * sizeh + i;
            const C rr = r[k], ss = s[k];
            const C rn = mkc<C>((double)rr.x + om * (double)ss.x, (double)rr.y + om * (double)ss.y);
            r[k] = rn;
            v += (double)rn.x * rn.x + (double)rn.y * rn.y;
        }
    }
    block_partial<1>(v, 0.0, ctl);
}

// p = p*gammag + r ; for active shifts k: x_k -= omega_k p_k ; p_k = gamma_k p_k + zeta_k r
// (inverter_multishift_full.c:138,145,158 = fermionic_utilities.c:315-378)
template <typename C>
__global__ __launch_bounds__(kBlock) void k_ms_update(const C* __restrict__ r, C* __restrict__ p, C* __restrict__ xs,
                                                      C* __restrict__ ps, long vstride, long n3, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n3) return;
    const SolverState* st = ctl.st;
    const double g = st->gammag;
    const C rr = r[i], pp = p[i];
    p[i] = mkc<C>((double)pp.x * g + (double)rr.x, (double)pp.y * g + (double)rr.y);
    const int cur = st->flag_cur, maxiter = st->maxiter;
    for (int k = 0; k < maxiter; k++) {
        if (st->flags[cur][k] != 1) continue;
        const double om = st->omegas[k], gk = st->gammas[k], zk = st->zeta_iii[k];
        C* xk = xs + k * vstride;
        C* pk = ps + k * vstride;
        const C pv = pk[i], xv = xk[i];
        xk[i] = mkc<C>((double)xv.x - om * (double)pv.x, (double)xv.y - om * (double)pv.y);
        pk[i] = mkc<C>(gk * (double)pv.x + zk * (double)rr.x, gk * (double)pv.y + zk * (double)rr.y);
    }
}

// ---------------------------------------------------------------------------
// CG-M with the shift updates applied once per window of W iterations.
//
// Within a window every shifted p_k and x_k is a linear combination of its
// value at the window start and of the window's base residuals r_0..r_{W-1}:
//   p_k(t) = B_k(t) p_k(0) + sum_{j<t} D_kj(t) r_j,   B(t+1) = gamma_k B(t),
//            D_kj(t+1) = gamma_k D_kj(t) (j < t), D_kt(t+1) = zeta_k
//   x_k(W) = x_k(0) + a_k p_k(0) + sum_j c_kj r_j,    a_k -= omega_k B(t),
//            c_kj -= omega_k D_kj(t)
// with the per-iteration scalars of the reference's recurrence
// (inverter_multishift_full.c:116-171; a frozen shift contributes nothing).
// The recurrence itself never reads x_k or p_k, so iteration counts and
// convergence flags are those of the per-iteration updates; the residuals
// are kept (W x 48 B per half-site) and one pass per window updates x_k and
// p_k: per iteration that is about 2 vectors of traffic instead of 4 per
// active shift.
constexpr int kMsWinMax = 64;
struct MsWin {
    int t;                                  // iterations in the current window
    int active[kMaxShifts];                 // shift moved in this window
    double a[kMaxShifts], B[kMaxShifts];
    double c[kMaxShifts][kMsWinMax], D[kMaxShifts][kMsWinMax];
};

// r_new = r_old + omega s ; lambda = |r_new|^2
template <typename C>
__global__ __launch_bounds__(kBlock) void k_msw_r(const C* __restrict__ r_old, C* __restrict__ r_new,
                                                  const C* __restrict__ s, int sizeh, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const double om = ctl.st->omega;
    double v = 0.0;
    if (i < sizeh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = (long)c * sizeh + i;
            const C rr = r_old[k], ss = s[k];
            const C rn = mkc<C>((double)rr.x + om * (double)ss.x, (double)rr.y + om * (double)ss.y);
            r_new[k] = rn;
            v += (double)rn.x * rn.x + (double)rn.y * rn.y;
        }
    }
    block_partial<1>(v, 0.0, ctl);
}

// p = gammag p + r; block 0 also advances the window coefficients of every
// active shift by this iteration's omega_k, gamma_k, zeta_k and the window
// position t.  With `slot`, r is also stored there as the window's residual
// t (the mixed CG-M keeps r in place; the host knows the window position,
// since windows start at fixed iterations).
template <typename C>
__global__ __launch_bounds__(kBlock) void k_msw_p(C* __restrict__ p, const C* __restrict__ r, long n3, MsWin* w,
                                                  SolverCtl ctl, C* __restrict__ slot) {
    if (!pred_pass(ctl)) return;
    const SolverState* st = ctl.st;
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n3) {
        const double g = st->gammag;
        const C rr = r[i], pp = p[i];
        p[i] = mkc<C>((double)pp.x * g + (double)rr.x, (double)pp.y * g + (double)rr.y);
        if (slot) slot[i] = rr;
    }
    if (blockIdx.x != 0) return;
    const int t = w->t;  // only block 0 reads and advances it
    const int cur = st->flag_cur, order = st->order;
    const int per = t + 2;  // j < t, the new column t, and the scalar pair
    for (int e = threadIdx.x; e < order * per; e += kBlock) {
        const int k = e / per, j = e - k * per;
        if (st->flags[cur][k] != 1) continue;
        const double om = st->omegas[k], gk = st->gammas[k];
        if (j < t) {
            w->c[k][j] -= om * w->D[k][j];
            w->D[k][j] *= gk;
        } else if (j == t) {
            w->D[k][t] = st->zeta_iii[k];
        } else {
            w->a[k] -= om * w->B[k];
            w->B[k] *= gk;
            w->active[k] = 1;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) w->t = t + 1;
}

// x_k += a_k p_k + sum_j c_kj r_j ; p_k = B_k p_k + sum_j D_kj r_j for the
// shifts that moved, kMswGroup shifts per pass over the residuals (eight: three
// passes for 19 shifts instead of five with four, each shift's sums unchanged).
// CX: the solutions' type (FP64 also when the directions and residuals C are
// FP32: the mixed CG-M adds each window straight into its FP64 solutions).
constexpr int kMswGroup = 8;
template <typename CX, typename C>
__global__ __launch_bounds__(kBlock) void k_msw_flush(CX* __restrict__ xs, C* __restrict__ ps,
                                                      const C* __restrict__ hist, long n3, int order,
                                                      const MsWin* __restrict__ w, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const int t = w->t;
    if (t == 0) return;
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n3) return;
    for (int k0 = 0; k0 < order; k0 += kMswGroup) {
        double2 ax[kMswGroup], ap[kMswGroup];
        bool on[kMswGroup];
        bool any = false;
#pragma unroll
        for (int q = 0; q < kMswGroup; q++) {
            const int k = k0 + q;
            on[q] = k < order && w->active[k];
            any = any || on[q];
            ax[q] = make_double2(0.0, 0.0);
            ap[q] = make_double2(0.0, 0.0);
            if (on[q]) {
                const CX xv = xs[k * n3 + i];
                const C pv = ps[k * n3 + i];
                const double a = w->a[k], b = w->B[k];
                ax[q] = make_double2((double)xv.x + a * (double)pv.x, (double)xv.y + a * (double)pv.y);
                ap[q] = make_double2(b * (double)pv.x, b * (double)pv.y);
            }
        }
        if (!any) continue;
        for (int j = 0; j < t; j++) {
            const C rv = hist[(long)j * n3 + i];
            const double rx = (double)rv.x, ry = (double)rv.y;
#pragma unroll
            for (int q = 0; q < kMswGroup; q++) {
                if (!on[q]) continue;
                const double c = w->c[k0 + q][j], d = w->D[k0 + q][j];
                ax[q].x += c * rx;
                ax[q].y += c * ry;
                ap[q].x += d * rx;
                ap[q].y += d * ry;
            }
        }
#pragma unroll
        for (int q = 0; q < kMswGroup; q++)
            if (on[q]) {
                xs[(k0 + q) * n3 + i] = mkc<CX>(ax[q].x, ax[q].y);
                ps[(k0 + q) * n3 + i] = mkc<C>(ap[q].x, ap[q].y);
            }
    }
}

// a fresh window: coefficients of the identity map
__global__ __launch_bounds__(kBlock) void k_msw_reset(MsWin* w, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    for (int e = threadIdx.x; e < kMaxShifts * kMsWinMax; e += kBlock) {
        const int k = e / kMsWinMax, j = e - k * kMsWinMax;
        w->c[k][j] = 0.0;
        w->D[k][j] = 0.0;
    }
    for (int k = threadIdx.x; k < kMaxShifts; k += kBlock) {
        w->a[k] = 0.0;
        w->B[k] = 1.0;
        w->active[k] = 0;
    }
    if (threadIdx.x == 0) w->t = 0;
}

// FP32 CG-M with FP64 reliable updates: r += omega s and, for the base
// system's solution (needed for the FP64 residual), x0_f -= omega p; |r|^2
__global__ __launch_bounds__(kBlock) void k_msx_r(float2* __restrict__ r, const float2* __restrict__ s,
                                                  const float2* __restrict__ p, float2* __restrict__ x0f, int sizeh,
                                                  SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const double om = ctl.st->omega;
    double v = 0.0;
    if (i < sizeh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = (long)c * sizeh + i;
            const float2 rr = r[k], ss = s[k], pp = p[k], xx = x0f[k];
            const float2 rn = make_float2((float)((double)rr.x + om * (double)ss.x), (float)((double)rr.y + om * (double)ss.y));
            r[k] = rn;
            x0f[k] = make_float2((float)((double)xx.x - om * (double)pp.x), (float)((double)xx.y - om * (double)pp.y));
            v += (double)rn.x * rn.x + (double)rn.y * rn.y;
        }
    }
    block_partial<1>(v, 0.0, ctl);
}

// x_k += x_k_f, x_k_f = 0 for the base system (slot 0) and the `order` shifts
// (slots 1..order): the FP32 accumulators are moved into the FP64 solutions
struct MsxFlush {
    double2* x64[kMaxShifts + 1];
    float2* x32[kMaxShifts + 1];
    int n;
};
__global__ __launch_bounds__(kBlock) void k_msx_flush(MsxFlush f, long n3, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n3) return;
    for (int k = 0; k < f.n; k++) {
        const float2 a = f.x32[k][i];
        const double2 x = f.x64[k][i];
        f.x64[k][i] = make_double2(x.x + (double)a.x, x.y + (double)a.y);
        f.x32[k][i] = make_float2(0.f, 0.f);
    }
}

// reliable update of the base residual: r64 = b - s64 (s64 = M x0 in FP64),
// r_f = (float) r64, |r_f|^2 of the stored values
__global__ __launch_bounds__(kBlock) void k_msx_resid(const double2* __restrict__ b, const double2* __restrict__ s,
                                                      float2* __restrict__ rf, int sizeh, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v = 0.0;
    if (i < sizeh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = (long)c * sizeh + i;
            const double2 bb = b[k], ss = s[k];
            const float2 r = make_float2((float)(bb.x - ss.x), (float)(bb.y - ss.y));
            rf[k] = r;
            v += (double)r.x * r.x + (double)r.y * r.y;
        }
    }
    block_partial<1>(v, 0.0, ctl);
}

// power method step (find_min_max.c:35-40): r = p / norm (the reference
// scales p in place and copies it to r; p is overwritten next by M r)
__global__ __launch_bounds__(kBlock) void k_pm_scale(const double2* __restrict__ p, double2* __restrict__ r, long n3,
                                                     SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n3) return;
    const double inorm = 1.0 / ctl.st->lambda;
    const double2 v = p[i];
    r[i] = make_double2(inorm * v.x, inorm * v.y);
}

// out = a0*in + sum_k a_k x_k   (inverter_multishift_full.c:254-282)
struct RecombineCoeffs {
    double a0;
    double a[kMaxShifts];
    int order;
};
__global__ __launch_bounds__(kBlock) void k_recombine(const double2* __restrict__ xs, const double2* __restrict__ in,
                                                      double2* __restrict__ out, long vstride, long n3,
                                                      RecombineCoeffs c) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n3) return;
    const double2 v = in[i];
    double ox = v.x * c.a0, oy = v.y * c.a0;
    for (int k = 0; k < c.order; k++) {
        const double2 x = xs[k * vstride + i];
        ox += c.a[k] * x.x;
        oy += c.a[k] * x.y;
    }
    out[i] = make_double2(ox, oy);
}

// ---- mixed precision ----
// r_f = (float)(b - s) ; p_f = r_f ; out_f = 0 ; delta = |r_f|^2   (inverter_mixedp.c:74-88)
__global__ __launch_bounds__(kBlock) void k_mx_init(const double2* __restrict__ b, const double2* __restrict__ s,
                                                    double2* __restrict__ r64, float2* __restrict__ rf,
                                                    float2* __restrict__ pf, float2* __restrict__ outf, int sizeh,
                                                    SolverCtl ctl) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v = 0.0;
    if (i < sizeh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = (long)c * sizeh + i;
            const double2 bb = b[k], ss = s[k];
            const double2 rr = make_double2(bb.x - ss.x, bb.y - ss.y);
            r64[k] = rr;
            const float2 f = make_float2((float)rr.x, (float)rr.y);
            rf[k] = f;
            pf[k] = f;
            outf[k] = make_float2(0.f, 0.f);
            v += (double)f.x * f.x + (double)f.y * f.y;
        }
    }
    block_partial<1>(v, 0.0, ctl);
}

// out_f += omega p_f; then either r_f -= omega s_f with lambda = |r_f|^2, or
// (reliable update) x += out_f, out_f = 0   (inverter_mixedp.c:107-134)
__global__ __launch_bounds__(kBlock) void k_mx_xr(float2* __restrict__ outf, const float2* __restrict__ pf,
                                                  float2* __restrict__ rf, const float2* __restrict__ sf,
                                                  double2* __restrict__ x, int sizeh, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const double om = ctl.st->omega;
    const int ru = ctl.st->ru;
    double v = 0.0;
    if (i < sizeh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = (long)c * sizeh + i;
            const float2 pp = pf[k], oo = outf[k];
            const float2 on = make_float2((float)((double)pp.x * om + (double)oo.x), (float)((double)pp.y * om + (double)oo.y));
            if (!ru) {
                outf[k] = on;
                const float2 ss = sf[k], rr = rf[k];
                const float2 rn = make_float2((float)((double)ss.x * (-om) + (double)rr.x),
                                              (float)((double)ss.y * (-om) + (double)rr.y));
                rf[k] = rn;
                v += (double)rn.x * rn.x + (double)rn.y * rn.y;
            } else {
                const double2 xx = x[k];
                x[k] = make_double2(xx.x + (double)on.x, xx.y + (double)on.y);
                outf[k] = make_float2(0.f, 0.f);
            }
        }
    }
    if (!ru) block_partial<1>(v, 0.0, ctl);
}

// reliable update tail: r = b - s (FP64), r_f = (float) r, lambda = |r_f|^2  (inverter_mixedp.c:121-124,137)
__global__ __launch_bounds__(kBlock) void k_mx_resid(const double2* __restrict__ b, const double2* __restrict__ s,
                                                     double2* __restrict__ r64, float2* __restrict__ rf, int sizeh,
                                                     SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v = 0.0;
    if (i < sizeh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = (long)c * sizeh + i;
            const double2 bb = b[k], ss = s[k];
            const double2 rr = make_double2(bb.x - ss.x, bb.y - ss.y);
            r64[k] = rr;
            const float2 f = make_float2((float)rr.x, (float)rr.y);
            rf[k] = f;
            v += (double)f.x * f.x + (double)f.y * f.y;
        }
    }
    block_partial<1>(v, 0.0, ctl);
}

// x += (double) out_f   (combine_add_in2_into_in1_mixed_precision, inverter_mixedp.c:24-34)
__global__ __launch_bounds__(kBlock) void k_add_f_to_d(double2* __restrict__ x, const float2* __restrict__ f, long n3) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n3) return;
    const double2 xx = x[i];
    const float2 ff = f[i];
    x[i] = make_double2(xx.x + (double)ff.x, xx.y + (double)ff.y);
}

// -------------------------------------------------------------- host driver
static inline hipStream_t S() { return ctx().stream; }
// iterations per HIP graph of the batched solves run inside the library
// (rational-solve refinement, chiral condensate); 0: direct launches
static int g_internal_ipg = 16;
static inline SolverState& H() { return *ctx().state_host; }

static SolverCtl ctl_of(int pred, int iter = 0) {
    SolverCtl c = plain_ctl();
    c.pred = pred;
    c.iter = iter;
    return c;
}

static bool check_launch(const char* what) { return hip_ok(hipGetLastError(), what); }

static int batch_next(int b) { return std::min(2 * b, 64); }

// Distance of a solver from its stopping test in log units of the residual,
// ln(residual / (res * safety)) (NaN when unknown): CG-type solvers from
// |r|^2, CG-M from the slowest active shift, |zeta_k|^2 |r|^2.
static double dist_cg(const SolverState& s) {
    if (!(s.lambda > 0.0) || !(s.source_norm > 0.0) || !(s.res > 0.0)) return NAN;
    return 0.5 * std::log(s.lambda / s.source_norm) - std::log(s.res * s.safety);
}
static double dist_ms(const SolverState& s) {
    const int act = s.flag_cur ^ 1;  // the flags the next iteration makes current
    double m = 0.0;
    for (int k = 0; k < s.order && k < kMaxShifts; k++)
        if (s.flags[act][k] == 1) m = std::max(m, s.delta * s.zeta_ii[k] * s.zeta_ii[k]);
    if (!(m > 0.0) || !(s.source_norm > 0.0) || !(s.res > 0.0)) return NAN;
    return 0.5 * std::log(m / s.source_norm) - std::log(s.res * s.safety);
}

// Queue iterations until the device state leaves kRunning.  Batches grow
// 4, 8, ..., 64; with a distance estimate, a batch is also cut to the
// iterations the convergence rate since the last batch predicts are left
// (plus a margin), so that few predicated no-op iterations follow the end.
template <typename Enq>
static bool run_batches(Enq enqueue, int first_batch = 4, double (*dist)(const SolverState&) = nullptr) {
    int batch = first_batch;
    int last_cg = -1;
    double prev_d = NAN;
    for (;;) {
        if (!pull_state()) return false;
        if (H().status != kRunning) return true;
        const int cg0 = H().cg;
        // every queued iteration of a running solver advances the counter on
        // the device; a stalled counter means the scalar epilogue did not run
        if (cg0 == last_cg) return fail("solver state did not advance (iteration %d)", cg0);
        int n = batch;
        if (dist) {
            const double d = dist(H());
            if (std::isfinite(d) && std::isfinite(prev_d) && last_cg >= 0 && cg0 > last_cg) {
                const double rate = (prev_d - d) / (double)(cg0 - last_cg);
                if (rate > 1e-9) {
                    const double left = d > 0.0 ? d / rate : 0.0;
                    n = std::min(n, std::max(2, (int)std::ceil(1.1 * left) + 1));
                }
            }
            prev_d = d;
        }
        last_cg = cg0;
        for (int j = 1; j <= n; j++)
            if (!enqueue(cg0 + j)) return false;
        batch = batch_next(batch);
    }
}

static void init_state(double res, double safety, int max_cg, int restart_every) {
    SolverState& s = H();
    std::memset(&s, 0, sizeof s);
    s.res = res;
    s.safety = safety;
    s.max_cg = max_cg;
    s.restart_every = restart_every;
    s.status = kRunning;
}

// CG of ker_invert_openacc on device pointers.  Returns status, *cg_out.
struct CgBufs {
    const void *u, *ph, *b;
    void *x, *r, *h, *s, *p;
};

static bool source_norm(const void* b, double* sn) { return reduce_to_host(0, b, nullptr, 0, sn, nullptr); }

int cg_solve(const CgBufs& B, double mass, double shift, double res, int max_cg, int restart_every, int* cg_out,
             double* true_res) {
    const long sizeh = ctx().lat.sizeh;
    const long n3 = 3 * sizeh;
    const int nbh = blocks_for(sizeh), nb3 = blocks_for(n3);
    const double m2 = mass * mass + shift;
    double sn = 0.0;
    if (!source_norm(B.b, &sn)) return INVERTER_FAILURE;
    init_state(res, 0.95, max_cg, restart_every);
    H().source_norm = sn;
    if (!push_state()) return INVERTER_FAILURE;

    auto restart = [&]() -> bool {
        // s = M x ; r = b - s ; p = r ; delta = |r|^2 ; cg_restarted = 0
        if (!launch_mdagm<double>(B.u, B.ph, B.x, B.h, B.s, m2, 1, kFinStore, ctl_of(kAlways))) return false;
        hipLaunchKernelGGL((k_cg_restart<kFinCgRestart>), dim3(nbh), dim3(kBlock), 0, S(), (const double2*)B.b,
                           (const double2*)B.s, (double2*)B.r, (double2*)B.p, (int)sizeh, ctl_of(kAlways));
        return check_launch("k_cg_restart") && launch_finalize(kFinCgRestart, ctl_of(kAlways), nbh);
    };
    auto iteration = [&](int j) -> bool {
        // s = M p ; alpha = (p,s) ; omega = delta/alpha
        if (!launch_mdagm<double>(B.u, B.ph, B.p, B.h, B.s, m2, 2, kFinCgAlpha, ctl_of(kIfRunning))) return false;
        hipLaunchKernelGGL((k_cg_xr<double2>), dim3(nbh), dim3(kBlock), 0, S(), (double2*)B.x, (double2*)B.r,
                           (const double2*)B.p, (const double2*)B.s, (int)sizeh, ctl_of(kIfRunning));
        if (!launch_finalize(kFinCgLambda, ctl_of(kIfRunning), nbh)) return false;
        hipLaunchKernelGGL((k_cg_p<double2>), dim3(nb3), dim3(kBlock), 0, S(), (double2*)B.p, (const double2*)B.r, n3,
                           ctl_of(kIfIter, j));
        return check_launch("cg iteration");
    };

    if (!restart()) return INVERTER_FAILURE;
    for (;;) {
        if (!run_batches(iteration, 4, dist_cg)) return INVERTER_FAILURE;
        const SolverState& s = H();
        if (s.status == kInnerExit && std::sqrt(s.lambda / s.source_norm) > res && s.cg < max_cg) {
            if (!restart()) return INVERTER_FAILURE;
            continue;
        }
        break;
    }
    const int cg = H().cg;
    // final true residual (inverter_full.c:114-130)
    if (!launch_mdagm<double>(B.u, B.ph, B.x, B.h, B.s, m2, 1, kFinStore, ctl_of(kAlways))) return INVERTER_FAILURE;
    hipLaunchKernelGGL((k_cg_restart<kFinStore>), dim3(nbh), dim3(kBlock), 0, S(), (const double2*)B.b,
                       (const double2*)B.s, (double2*)B.h, (double2*)nullptr, (int)sizeh, ctl_of(kAlways));
    if (!launch_finalize(kFinStore, ctl_of(kAlways), nbh)) return INVERTER_FAILURE;
    if (!check_launch("final residual") || !pull_state()) return INVERTER_FAILURE;
    const double current_res = H().sum / sn;
    if (verbosity_lv > 1)
        printf("Terminated invert after   %d    iterations[res/stop_res=  %e , stop_res=%e ]\n", cg,
               std::sqrt(current_res) / res, res);
    if (cg == max_cg) printf("WARNING: maximum number of iterations reached in invert\n");
    *cg_out = cg;
    if (true_res) *true_res = std::sqrt(current_res);
    return (std::sqrt(current_res) <= res) ? INVERTER_SUCCESS : INVERTER_FAILURE;
}

// Pieces reused by the flexible CG of the learned preconditioner (ptc.hip).
bool launch_cg_xr(void* x, void* r, const void* p, const void* s, const SolverCtl& ctl) {
    const long sizeh = ctx().lat.sizeh;
    hipLaunchKernelGGL((k_cg_xr<double2>), dim3(blocks_for(sizeh)), dim3(kBlock), 0, S(), (double2*)x, (double2*)r,
                       (const double2*)p, (const double2*)s, (int)sizeh, ctl);
    return check_launch("k_cg_xr");
}
bool launch_cg_p(void* p, const void* r, const SolverCtl& ctl) {
    const long n3 = 3 * ctx().lat.sizeh;
    hipLaunchKernelGGL((k_cg_p<double2>), dim3(blocks_for(n3)), dim3(kBlock), 0, S(), (double2*)p, (const double2*)r,
                       n3, ctl);
    return check_launch("k_cg_p");
}
// The batch driver for solvers outside this file (flexible CG): iterations
// queued through `enqueue`, batches cut by the |r|^2 convergence rate.
bool run_cg_batches(const std::function<bool(int)>& enqueue, int first_batch) {
    return run_batches([&](int j) { return enqueue(j); }, first_batch, dist_cg);
}
// FP32 forms for the FP32-inner flexible CG; x += (double) e
bool launch_cg_xr_f(void* x, void* r, const void* p, const void* s, const SolverCtl& ctl) {
    const long sizeh = ctx().lat.sizeh;
    hipLaunchKernelGGL((k_cg_xr<float2>), dim3(blocks_for(sizeh)), dim3(kBlock), 0, S(), (float2*)x, (float2*)r,
                       (const float2*)p, (const float2*)s, (int)sizeh, ctl);
    return check_launch("k_cg_xr<float2>");
}
bool launch_cg_p_f(void* p, const void* r, const SolverCtl& ctl) {
    const long n3 = 3 * ctx().lat.sizeh;
    hipLaunchKernelGGL((k_cg_p<float2>), dim3(blocks_for(n3)), dim3(kBlock), 0, S(), (float2*)p, (const float2*)r, n3,
                       ctl);
    return check_launch("k_cg_p<float2>");
}
bool launch_add_f_to_d(void* x, const void* e) {
    const long n3 = 3 * ctx().lat.sizeh;
    hipLaunchKernelGGL(k_add_f_to_d, dim3(blocks_for(n3)), dim3(kBlock), 0, S(), (double2*)x, (const float2*)e, n3);
    return check_launch("k_add_f_to_d");
}
int cg_solve_from(const void* u, const void* ph, const void* b, void* x, void* r, void* h, void* s, void* p,
                  double mass, double res, int max_cg, int* cg_out, double* true_res) {
    CgBufs B;
    B.u = u;
    B.ph = ph;
    B.b = b;
    B.x = x;
    B.r = r;
    B.h = h;
    B.s = s;
    B.p = p;
    return cg_solve(B, mass, 0.0, res, max_cg, inverter_tricks.restartingEvery, cg_out, true_res);
}

// Mixed precision CG (inverter_mixed_precision) on device pointers.
struct MxBufs {
    const void *u, *u_f, *ph, *ph_f, *b;
    void *x, *r, *h, *s;              // FP64 scratch (ip.loc_r/h/s)
    void *rf, *hf, *sf, *pf, *outf;   // FP32 scratch (ip.loc_*_f, ip.out_f)
};

int mixed_solve(const MxBufs& B, double mass, double shift, double res, int max_cg, double mixed_delta,
                int* cg_out, int* ru_count, double* true_res) {
    const long sizeh = ctx().lat.sizeh;
    const long n3 = 3 * sizeh;
    const int nbh = blocks_for(sizeh), nb3 = blocks_for(n3);
    const double m2 = mass * mass + shift;
    const double m2f = mass * mass + (double)(float)shift;  // sp_ code passes `float shift`
    double sn = 0.0;
    if (!source_norm(B.b, &sn)) return INVERTER_FAILURE;
    init_state(res, 0.9, max_cg, 0);
    H().source_norm = sn;
    H().mixed_delta = mixed_delta;
    if (!push_state()) return INVERTER_FAILURE;

    // r = b - M x (FP64) ; r_f = (float) r ; p_f = r_f ; out_f = 0 ; delta = |r_f|^2
    if (!launch_mdagm<double>(B.u, B.ph, B.x, B.h, B.s, m2, 1, kFinStore, ctl_of(kAlways))) return INVERTER_FAILURE;
    hipLaunchKernelGGL(k_mx_init, dim3(nbh), dim3(kBlock), 0, S(), (const double2*)B.b, (const double2*)B.s,
                       (double2*)B.r, (float2*)B.rf, (float2*)B.pf, (float2*)B.outf, (int)sizeh, ctl_of(kAlways));
    if (!check_launch("k_mx_init") || !launch_finalize(kFinCgRestart, ctl_of(kAlways), nbh)) return INVERTER_FAILURE;

    auto iteration = [&](int j) -> bool {
        if (!launch_mdagm<float>(B.u_f, B.ph_f, B.pf, B.hf, B.sf, m2f, 2, kFinMxAlpha, ctl_of(kIfRunning)))
            return false;
        hipLaunchKernelGGL(k_mx_xr, dim3(nbh), dim3(kBlock), 0, S(), (float2*)B.outf, (const float2*)B.pf,
                           (float2*)B.rf, (const float2*)B.sf, (double2*)B.x, (int)sizeh, ctl_of(kIfRunning));
        if (!launch_finalize(kFinMxLambda, ctl_of(kIfRunningNoRU), nbh)) return false;
        // reliable update: s = M x (FP64), r = b - s, r_f = (float) r
        if (!launch_mdagm<double>(B.u, B.ph, B.x, B.h, B.s, m2, 1, kFinStore, ctl_of(kIfRunningRU))) return false;
        hipLaunchKernelGGL(k_mx_resid, dim3(nbh), dim3(kBlock), 0, S(), (const double2*)B.b, (const double2*)B.s,
                           (double2*)B.r, (float2*)B.rf, (int)sizeh, ctl_of(kIfRunningRU));
        if (!launch_finalize(kFinMxLambda, ctl_of(kIfRunningRU), nbh)) return false;
        hipLaunchKernelGGL((k_cg_p<float2>), dim3(nb3), dim3(kBlock), 0, S(), (float2*)B.pf, (const float2*)B.rf, n3,
                           ctl_of(kIfIter, j));
        return check_launch("mixed iteration");
    };
    if (!run_batches(iteration, 4, dist_cg)) return INVERTER_FAILURE;
    const int cg = H().cg;
    if (ru_count) *ru_count = H().ru_count;
    hipLaunchKernelGGL(k_add_f_to_d, dim3(nb3), dim3(kBlock), 0, S(), (double2*)B.x, (const float2*)B.outf, n3);
    if (!launch_mdagm<double>(B.u, B.ph, B.x, B.h, B.s, m2, 1, kFinStore, ctl_of(kAlways))) return INVERTER_FAILURE;
    hipLaunchKernelGGL((k_cg_restart<kFinStore>), dim3(nbh), dim3(kBlock), 0, S(), (const double2*)B.b,
                       (const double2*)B.s, (double2*)B.h, (double2*)nullptr, (int)sizeh, ctl_of(kAlways));
    if (!launch_finalize(kFinStore, ctl_of(kAlways), nbh)) return INVERTER_FAILURE;
    if (!check_launch("final residual") || !pull_state()) return INVERTER_FAILURE;
    const double giustoono = H().sum / sn;
    if (verbosity_lv > 1)
        printf("Terminated invert after   %d    iterations[res/stop_res=  %e , stop_res=%e ]\n", cg,
               std::sqrt(giustoono) / res, res);
    if (cg == max_cg) printf("WARNING: maximum number of iterations reached in invert\n");
    *cg_out = cg;
    if (true_res) *true_res = std::sqrt(giustoono);
    return (std::sqrt(giustoono) <= res) ? INVERTER_SUCCESS : INVERTER_FAILURE;
}

// ------------------------------------------- per-shift verification (FP64)
// The check after the CG-M loop (inverter_multishift_full.c:209-225):
// |b - (M + sigma_k) x_k|^2 for every shift.  The shifts share the operator,
// so chunks of up to kVerShifts solutions go through the two hops together:
// each link is loaded once per half-site and applied to every solution of
// the chunk (hop_mr), instead of one M application per shift.
constexpr int kVerShifts = 8;
struct VerArgs {
    const double2* u[8];
    const double* ph[8];
    const double2* x[kVerShifts];
    double m2[kVerShifts];
    int n;             // solutions in this chunk
    double2* h;        // kVerShifts D_oe x_k
    const double2* b;
    double* part;      // kVerShifts x gridDim.x block sums of |r_k|^2
    SiteGeom g;
};

// EO = 1: h_k = D_oe x_k ; EO = 0: r_k = b - (m2_k x_k - D_eo h_k), block sums of |r_k|^2
template <int EO>
__global__ __launch_bounds__(kBlock) void k_ms_verify(VerArgs a) {
    const int blk = xcd_block_id();
    const int idx = blk * kBlock + (int)threadIdx.x;
    const long n3 = 3L * a.g.sizeh;
    bool act[kVerShifts];
#pragma unroll
    for (int r = 0; r < kVerShifts; r++) act[r] = r < a.n;
    double dot[kVerShifts];
#pragma unroll
    for (int r = 0; r < kVerShifts; r++) dot[r] = 0.0;
    if (idx < a.g.sizeh) {
        const double2* in[kVerShifts];
#pragma unroll
        for (int r = 0; r < kVerShifts; r++) in[r] = EO == 1 ? a.x[act[r] ? r : 0] : a.h + (act[r] ? r : 0) * n3;
        double2 o[kVerShifts][3];
#pragma unroll
        for (int r = 0; r < kVerShifts; r++) o[r][0] = o[r][1] = o[r][2] = make_double2(0.0, 0.0);
        const Nbrs nb = site_neighbours<EO>(a.g, idx);
        constexpr int F = EO == 0 ? 0 : 1, B = 1 - F;
#pragma unroll
        for (int mu = 0; mu < 4; mu++)
            hop_mr<double, true, kVerShifts>(a.u[2 * mu + B], a.ph[2 * mu + B], in, act, a.g.sizeh, nb.im[mu], nb.im[mu], o);
#pragma unroll
        for (int mu = 0; mu < 4; mu++)
            hop_mr<double, false, kVerShifts>(a.u[2 * mu + F], a.ph[2 * mu + F], in, act, a.g.sizeh, idx, nb.ip[mu], o);
#pragma unroll
        for (int r = 0; r < kVerShifts; r++) {
            if (!act[r]) continue;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double2 oc = make_double2(0.5 * o[r][c].x, 0.5 * o[r][c].y);
                const long k = (long)c * a.g.sizeh + idx;
                if constexpr (EO == 1) {
                    a.h[r * n3 + k] = oc;
                } else {
                    const double2 xx = a.x[r][k], bb = a.b[k];
                    const double2 sv = make_double2(xx.x * a.m2[r] - oc.x, xx.y * a.m2[r] - oc.y);
                    const double2 rr = make_double2(bb.x - sv.x, bb.y - sv.y);
                    dot[r] += rr.x * rr.x + rr.y * rr.y;
                }
            }
        }
    }
    if constexpr (EO == 0) {
        __shared__ double lds[kVerShifts][kBlock / 64];
#pragma unroll
        for (int r = 0; r < kVerShifts; r++) {
            if (!act[r]) continue;
            const double t = block_sum<kBlock>(dot[r], lds[r]);
            if (threadIdx.x == 0) a.part[(long)r * gridDim.x + blk] = t;
        }
    }
}

// sum of the block partials of solution blockIdx.x, in block order
__global__ __launch_bounds__(kBlock) void k_ms_verify_sum(const double* __restrict__ part, int nb, double* out) {
    __shared__ double lds[kBlock / 64];
    double v = 0.0;
    for (int i = threadIdx.x; i < nb; i += kBlock) v += part[(long)blockIdx.x * nb + i];
    const double t = block_sum<kBlock>(v, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

static void* g_ver_buf = nullptr;
static size_t g_ver_bytes = 0;

static VerArgs ver_args(const void* u, const void* ph) {
    const Lattice& L = ctx().lat;
    VerArgs a;
    for (int l = 0; l < 8; l++) {
        a.u[l] = (const double2*)((const char*)u + l * L.su3_bytes());
        a.ph[l] = (const double*)((const char*)ph + l * L.dsoa_bytes());
    }
    a.g.nd0 = L.nd[0];
    a.g.nd0h = L.nd[0] / 2;
    a.g.nd1 = L.nd[1];
    a.g.nd2 = L.nd[2];
    a.g.nd3 = L.nd[3];
    a.g.sizeh = (int)L.sizeh;
    a.b = nullptr;
    a.part = nullptr;
    return a;
}

// hs_k = D_oe xs_k for n FP64 even fields xs_k = xs + k * vb (the fermion
// force's odd parts), chunks of kVerShifts per launch
bool launch_doe_multi(const void* u, const void* ph, const void* xs, void* hs, int n) {
    const Lattice& L = ctx().lat;
    const size_t vb = L.vec3_bytes();
    VerArgs a = ver_args(u, ph);
    for (int k0 = 0; k0 < n; k0 += kVerShifts) {
        a.n = std::min(kVerShifts, n - k0);
        for (int r = 0; r < kVerShifts; r++) {
            a.x[r] = (const double2*)((const char*)xs + (k0 + (r < a.n ? r : 0)) * vb);
            a.m2[r] = 0.0;
        }
        a.h = (double2*)((char*)hs + k0 * vb);
        hipLaunchKernelGGL(k_ms_verify<1>, dim3(blocks_for(L.sizeh)), dim3(kBlock), 0, S(), a);
    }
    return check_launch("k_ms_verify<1>");
}

// |b - (M + shifts_k) x_k|^2 / sn for the FP64 solutions x_k = out + k * vb
static bool verify_shifts_f64(const void* u, const void* ph, const void* b, const void* out, int order, double m2,
                              const double* shifts, double sn, double* g) {
    const Lattice& L = ctx().lat;
    const long sizeh = L.sizeh;
    const size_t vb = L.vec3_bytes();
    const int nb = blocks_for(sizeh);
    const size_t need = kVerShifts * vb + ((size_t)kVerShifts * nb + kVerShifts) * sizeof(double) + 256;
    if (g_ver_bytes < need) {
        if (g_ver_buf) (void)hipFree(g_ver_buf);
        g_ver_buf = nullptr;
        g_ver_bytes = 0;
        if (!hip_ok(hipMalloc(&g_ver_buf, need), "multishift verification scratch")) return false;
        g_ver_bytes = need;
    }
    VerArgs a = ver_args(u, ph);
    a.h = (double2*)g_ver_buf;
    a.part = (double*)((char*)g_ver_buf + kVerShifts * vb);
    double* sums = a.part + (size_t)kVerShifts * nb;
    a.b = (const double2*)b;
    for (int k0 = 0; k0 < order; k0 += kVerShifts) {
        a.n = std::min(kVerShifts, order - k0);
        for (int r = 0; r < kVerShifts; r++) {
            const int k = k0 + (r < a.n ? r : 0);
            a.x[r] = (const double2*)((const char*)out + k * vb);
            a.m2[r] = m2 + shifts[k];
        }
        hipLaunchKernelGGL(k_ms_verify<1>, dim3(nb), dim3(kBlock), 0, S(), a);
        hipLaunchKernelGGL(k_ms_verify<0>, dim3(nb), dim3(kBlock), 0, S(), a);
        hipLaunchKernelGGL(k_ms_verify_sum, dim3(a.n), dim3(kBlock), 0, S(), (const double*)a.part, nb, sums);
        double h_sums[kVerShifts];
        if (!check_launch("multishift verification") ||
            !hip_ok(hipMemcpyAsync(h_sums, sums, sizeof(double) * a.n, hipMemcpyDeviceToHost, S()), "verification") ||
            !hip_ok(hipStreamSynchronize(S()), "verification"))
            return false;
        for (int r = 0; r < a.n; r++) g[k0 + r] = h_sums[r] / sn;
    }
    return true;
}

// CG-M (multishift_invert) on device pointers; R = double or float vectors
// (the float variant is the sp_ multishift_invert_f: float data, double scalars).
struct MsBufs {
    const void *u, *ph, *b;
    void *out, *r, *h, *s, *p, *shiftferm;
};

static void* g_msw_hist = nullptr;
static size_t g_msw_hist_bytes = 0;
static MsWin* g_msw = nullptr;
static int g_msw_forced = 0;  // 0 automatic, 1 per-iteration updates, 2..64 fixed

// window length: as many residuals as fit in a quarter of the free memory,
// at most 64 (fewer than 8: per-iteration updates)
static int msw_window(size_t vb) {
    if (g_msw_forced == 1) return 0;
    if (g_msw_forced >= 2) return g_msw_forced;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        free_b = 0;
    }
    const size_t have = g_msw_hist_bytes + free_b / 4;
    size_t w = vb ? have / vb : 0;
    if (w > (size_t)kMsWinMax) w = kMsWinMax;
    return w >= 8 ? (int)w : 0;
}

template <typename R>
int multishift_solve(const MsBufs& B, double mass, int order, const double* shifts, double residuo, int max_cg,
                     int* cg_out, double* shift_res) {
    using C = typename std::conditional<sizeof(R) == 8, double2, float2>::type;
    const Lattice& L = ctx().lat;
    const long sizeh = L.sizeh;
    const long n3 = 3 * sizeh;
    const long vstride = n3;  // in complex elements
    const size_t vb = sizeof(R) == 8 ? L.vec3_bytes() : L.vec3f_bytes();
    const int nbh = blocks_for(sizeh), nb3 = blocks_for(n3);
    const double m2 = mass * mass;
    if (order < 0 || order > kMaxShifts) return fail("multishift: order %d out of range", order) ? 0 : INVERTER_FAILURE;

    // x_k = 0 ; r = p = b ; p_k = b
    for (int k = 0; k < order; k++) {
        if (!hip_ok(hipMemsetAsync((char*)B.out + k * vb, 0, vb, S()), "zero x_k")) return INVERTER_FAILURE;
        if (!hip_ok(hipMemcpyAsync((char*)B.shiftferm + k * vb, B.b, vb, hipMemcpyDeviceToDevice, S()), "p_k = b"))
            return INVERTER_FAILURE;
    }
    if (!hip_ok(hipMemcpyAsync(B.r, B.b, vb, hipMemcpyDeviceToDevice, S()), "r = b") ||
        !hip_ok(hipMemcpyAsync(B.p, B.b, vb, hipMemcpyDeviceToDevice, S()), "p = b"))
        return INVERTER_FAILURE;
    double sn = 0.0;
    bool ok = sizeof(R) == 8 ? reduce_to_host(0, B.b, nullptr, 0, &sn, nullptr)
                             : reduce_to_host(0, B.b, nullptr, 1, &sn, nullptr);
    if (!ok) return INVERTER_FAILURE;
    init_state(residuo, 0.95, max_cg, 0);
    SolverState& s = H();
    s.delta = sn;
    s.source_norm = sn;
    s.omega = 1.0;
    s.gammag = 0.0;
    s.order = order;
    s.flag_cur = 0;
    for (int k = 0; k < order; k++) {
        s.flags[1][k] = 1;  // becomes current at the first iteration
        s.zeta_i[k] = 1.0;
        s.zeta_ii[k] = 1.0;
        s.gammas[k] = 0.0;
        s.shifts[k] = shifts[k];
    }
    s.maxiter_next = order;
    if (order == 0) s.status = kDone;  // the reference's loop runs once; nothing to update without shifts
    if (!push_state()) return INVERTER_FAILURE;

    // residual window (falls back to per-iteration shift updates when fewer
    // than 8 residuals fit)
    const int W = msw_window(vb);
    char* hist = nullptr;
    if (W >= 2) {
        hist = (char*)g_msw_hist;
        if (g_msw_hist_bytes < (size_t)W * vb) {
            if (g_msw_hist) (void)hipFree(g_msw_hist);
            g_msw_hist = nullptr;
            g_msw_hist_bytes = 0;
            if (!hip_ok(hipMalloc(&g_msw_hist, (size_t)W * vb), "multishift residual window")) return INVERTER_FAILURE;
            g_msw_hist_bytes = (size_t)W * vb;
        }
        hist = (char*)g_msw_hist;
        if (!g_msw && !hip_ok(hipMalloc(&g_msw, sizeof(MsWin)), "multishift window")) return INVERTER_FAILURE;
        hipLaunchKernelGGL(k_msw_reset, dim3(1), dim3(kBlock), 0, S(), g_msw, ctl_of(kAlways));
    }
    auto slot = [&](int j) { return (C*)(hist + (size_t)((j - 1) % W) * vb); };  // residual of iteration j (1-based)
    auto iteration = [&](int j) -> bool {
        if (!launch_mdagm<R>(B.u, B.ph, B.p, B.h, B.s, m2, 2, kFinMsAlpha, ctl_of(kIfRunning))) return false;
        if (!hist) {
            hipLaunchKernelGGL((k_ms_r<C>), dim3(nbh), dim3(kBlock), 0, S(), (C*)B.r, (const C*)B.s, (int)sizeh,
                               ctl_of(kIfRunning));
            if (!launch_finalize(kFinMsLambda, ctl_of(kIfRunning), nbh)) return false;
            hipLaunchKernelGGL((k_ms_update<C>), dim3(nb3), dim3(kBlock), 0, S(), (const C*)B.r, (C*)B.p,
                               (C*)B.out, (C*)B.shiftferm, vstride, n3, ctl_of(kIfIter, j));
            return check_launch("multishift iteration");
        }
        const C* r_old = j == 1 ? (const C*)B.r : slot(j - 1);
        C* r_new = slot(j);
        hipLaunchKernelGGL((k_msw_r<C>), dim3(nbh), dim3(kBlock), 0, S(), r_old, r_new, (const C*)B.s, (int)sizeh,
                           ctl_of(kIfRunning));
        if (!launch_finalize(kFinMsLambda, ctl_of(kIfRunning), nbh)) return false;
        hipLaunchKernelGGL((k_msw_p<C>), dim3(nb3), dim3(kBlock), 0, S(), (C*)B.p, (const C*)r_new, n3, g_msw,
                           ctl_of(kIfIter, j), (C*)nullptr);
        if (j % W == 0) {
            // only when iteration j ran: a window cut short is flushed once, at the end
            hipLaunchKernelGGL((k_msw_flush<C, C>), dim3(nb3), dim3(kBlock), 0, S(), (C*)B.out, (C*)B.shiftferm,
                               (const C*)hist, n3, order, (const MsWin*)g_msw, ctl_of(kIfIter, j));
            hipLaunchKernelGGL(k_msw_reset, dim3(1), dim3(kBlock), 0, S(), g_msw, ctl_of(kIfIter, j));
        }
        return check_launch("multishift iteration");
    };
    if (!run_batches(iteration, 4, dist_ms)) return INVERTER_FAILURE;
    const int cg = H().cg;
    if (hist) {
        // the partial last window, and the final residual where the per-iteration
        // version leaves it
        hipLaunchKernelGGL((k_msw_flush<C, C>), dim3(nb3), dim3(kBlock), 0, S(), (C*)B.out, (C*)B.shiftferm,
                           (const C*)hist, n3, order, (const MsWin*)g_msw, ctl_of(kAlways));
        if (cg > 0 && !hip_ok(hipMemcpyAsync(B.r, slot(cg), vb, hipMemcpyDeviceToDevice, S()), "final residual"))
            return INVERTER_FAILURE;
        if (!check_launch("multishift window flush")) return INVERTER_FAILURE;
    }
    multishift_invert_iterations += cg;
    if (cg == max_cg) printf("WARNING: maximum number of iterations reached in invert\n");

    // per-shift verification with (M + sigma_k) (inverter_multishift_full.c:212-229)
    int check = 1;
    if constexpr (sizeof(R) == 8) {
        double g[kMaxShifts];
        if (order > 0 && !verify_shifts_f64(B.u, B.ph, B.b, B.out, order, m2, shifts, sn, g)) return INVERTER_FAILURE;
        for (int k = 0; k < order; k++) {
            if (shift_res) shift_res[k] = std::sqrt(g[k]);
            check *= (g[k] <= 1) ? 1 : 0;
        }
    }
    for (int k = 0; sizeof(R) != 8 && k < order; k++) {
        const double m2k = sizeof(R) == 8 ? m2 + shifts[k] : m2 + (double)(float)shifts[k];
        const void* xk = (const char*)B.out + k * vb;
        if (!launch_mdagm<R>(B.u, B.ph, xk, B.h, B.s, m2k, 1, kFinStore, ctl_of(kAlways))) return INVERTER_FAILURE;
        double g = 0.0;
        if (sizeof(R) == 8) {
            hipLaunchKernelGGL((k_cg_restart<kFinStore>), dim3(nbh), dim3(kBlock), 0, S(), (const double2*)B.b,
                               (const double2*)B.s, (double2*)B.h, (double2*)nullptr, (int)sizeh, ctl_of(kAlways));
            if (!launch_finalize(kFinStore, ctl_of(kAlways), nbh)) return INVERTER_FAILURE;
            if (!check_launch("verify") || !pull_state()) return INVERTER_FAILURE;
            g = H().sum / sn;
        } else {
            if (!launch_axpy<float2>(B.s, -1.0, B.b, B.h, ctl_of(kAlways))) return INVERTER_FAILURE;
            double nn = 0.0;
            if (!reduce_to_host(0, B.h, nullptr, 1, &nn, nullptr)) return INVERTER_FAILURE;
            g = nn / sn;
        }
        if (shift_res) shift_res[k] = std::sqrt(g);
        check *= (g <= 1) ? 1 : 0;
    }
    if (!hip_ok(hipMemcpyAsync(B.p, (const char*)B.out + (order > 0 ? (order - 1) : 0) * vb, vb,
                               hipMemcpyDeviceToDevice, S()),
                "loc_p"))
        return INVERTER_FAILURE;
    if (verbosity_lv > 0)
        printf("Terminated multishift_invert ( target res = %1.1e,source_norm = %1.1e )\tCG count %d\n", residuo, sn,
               cg);
    *cg_out = cg;
    if (!hip_ok(hipStreamSynchronize(S()), "synchronize")) return INVERTER_FAILURE;
    return check == 1 ? INVERTER_SUCCESS : INVERTER_FAILURE;
}

template int multishift_solve<double>(const MsBufs&, double, int, const double*, double, int, int*, double*);
template int multishift_solve<float>(const MsBufs&, double, int, const double*, double, int, int*, double*);

// ------------------------------------------- several CG-Ms on one operator
// Rational solves of several sources whose operators differ only by the mass
// (the flavours and pseudofermions of an RHMC force on one smeared gauge
// field, when their U(1) phases agree) advance together: each iteration's
// two hops run once for all of them, every link applied to each system's
// spinor (hop_mr, the single-system hop's arithmetic, with k_dslash's tile
// order and reduction slots), and every other kernel of the iteration runs
// per system on its own device state, window and partials.  Each system's
// iterations, solutions and residuals are those of multishift_solve.
constexpr int kMsMulti = 4;
struct MsMultiHop {
    const double2* u[8];
    const double* ph[8];
    const double2* in[kMsMulti];
    double2* out[kMsMulti];
    const double2* x[kMsMulti];   // EO = 0: out = m2 x - D_eo in, <x, out> partials
    double m2[kMsMulti];
    SolverState* st[kMsMulti];
    double* part[kMsMulti];
    int n, pred, tmajor, rev;
    SiteGeom g;
};

template <int EO, int NR>
__global__ __launch_bounds__(kBlock) void k_ms_multi_hop(MsMultiHop a) {
    static_assert(NR <= kMsMulti, "systems per launch");
    bool act[NR];
    int first = -1;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        SolverCtl c;
        c.st = a.st[r < a.n ? r : 0];
        c.partials = a.part[r < a.n ? r : 0];
        c.pred = a.pred;
        act[r] = r < a.n && pred_pass(c);
        if (act[r] && first < 0) first = r;
    }
    if (first < 0) return;
    // k_dslash's workgroup -> tile map (and so its reduction slots)
    const int blk = xcd_block_id_dir(a.rev);
    int idx;
    if (a.tmajor) {
        const int s3 = a.g.sizeh / a.g.nd3;
        const int tile = blk / a.g.nd3, tt = blk - tile * a.g.nd3;
        idx = tt * s3 + tile * kBlock + (int)threadIdx.x;
    } else {
        idx = blk * kBlock + (int)threadIdx.x;
    }
    double dot[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) dot[r] = 0.0;
    if (idx < a.g.sizeh) {
        const double2* in[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) in[r] = a.in[act[r] ? r : first];
        double2 o[NR][3];
#pragma unroll
        for (int r = 0; r < NR; r++) o[r][0] = o[r][1] = o[r][2] = make_double2(0.0, 0.0);
        const Nbrs nb = site_neighbours<EO>(a.g, idx);
        constexpr int F = EO == 0 ? 0 : 1, B = 1 - F;
#pragma unroll
        for (int mu = 0; mu < 4; mu++)
            hop_mr<double, true, NR>(a.u[2 * mu + B], a.ph[2 * mu + B], in, act, a.g.sizeh, nb.im[mu],
                                           nb.im[mu], o);
#pragma unroll
        for (int mu = 0; mu < 4; mu++)
            hop_mr<double, false, NR>(a.u[2 * mu + F], a.ph[2 * mu + F], in, act, a.g.sizeh, idx, nb.ip[mu],
                                            o);
        const long sh = a.g.sizeh;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (!act[r]) continue;
            double2 o0 = o[r][0], o1 = o[r][1], o2 = o[r][2];
            const double half = 0.5;
            o0.x *= half; o0.y *= half; o1.x *= half; o1.y *= half; o2.x *= half; o2.y *= half;
            if constexpr (EO == 1) {
                st_nt(a.out[r] + idx, o0);
                st_nt(a.out[r] + sh + idx, o1);
                st_nt(a.out[r] + 2 * sh + idx, o2);
            } else {
                const double2* x = a.x[r];
                const double2 x0 = x[idx], x1 = x[sh + idx], x2 = x[2 * sh + idx];
                const double m2 = a.m2[r];
                const double2 s0 = make_double2(x0.x * m2 - o0.x, x0.y * m2 - o0.y);
                const double2 s1 = make_double2(x1.x * m2 - o1.x, x1.y * m2 - o1.y);
                const double2 s2 = make_double2(x2.x * m2 - o2.x, x2.y * m2 - o2.y);
                st_nt(a.out[r] + idx, s0);
                st_nt(a.out[r] + sh + idx, s1);
                st_nt(a.out[r] + 2 * sh + idx, s2);
                dot[r] = re_dot(x0, s0) + re_dot(x1, s1) + re_dot(x2, s2);
            }
        }
    }
    if constexpr (EO == 0) {
        // block_partial_at's sums, one LDS row per system (no wave may overwrite
        // a wave sum of another system before thread 0 has read it)
        __shared__ double lds[NR][kBlock / 64];
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            double v = dot[r];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) lds[r][wid] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll
            for (int r = 0; r < NR; r++) {
                if (!act[r]) continue;
                double t = 0.0;
#pragma unroll
                for (int w = 0; w < kBlock / 64; w++) t += lds[r][w];
                a.part[r][2 * blk] = t;
            }
        }
    }
}

// two systems (the common case: two flavours) with the two-system form
template <int EO>
static void launch_multi_hop(const MsMultiHop& a, int nbh) {
    if (a.n <= 2)
        hipLaunchKernelGGL((k_ms_multi_hop<EO, 2>), dim3(nbh), dim3(kBlock), 0, S(), a);
    else
        hipLaunchKernelGGL((k_ms_multi_hop<EO, kMsMulti>), dim3(nbh), dim3(kBlock), 0, S(), a);
}

struct MsSys {
    const void* b;
    void *out, *shiftferm, *r, *h, *s, *p;
    double mass;
    int order;
    const double* shifts;
    double residuo;
    int max_cg;
    int cg = 0;          // out: CG-M iterations
    int status = INVERTER_FAILURE;
};

static SolverState* g_mm_st = nullptr;       // device states of the systems
static SolverState* g_mm_st_host = nullptr;  // pinned mirror
static double* g_mm_part = nullptr;          // kMsMulti x 2 nbh partials
static size_t g_mm_part_n = 0;
static MsWin* g_mm_win = nullptr;            // kMsMulti windows
static void* g_mm_hist = nullptr;            // kMsMulti x W residuals
static size_t g_mm_hist_bytes = 0;

// multishift_solve<double> for n <= kMsMulti systems on the operator (u, ph);
// false (nothing changed) when the window memory is not there, so the caller
// solves them one by one
static bool multishift_solve_multi(const void* u, const void* ph, MsSys* sy, int n) {
    const Lattice& L = ctx().lat;
    const long sizeh = L.sizeh, n3 = 3 * sizeh;
    const size_t vb = L.vec3_bytes();
    const int nbh = blocks_for(sizeh), nb3 = blocks_for(n3);
    if (n < 1 || n > kMsMulti) return false;
    for (int k = 0; k < n; k++)
        if (sy[k].order < 0 || sy[k].order > kMaxShifts) return false;
    const int W = msw_window(vb);
    if (W < 2) return false;
    // device resources
    if (!g_mm_st) {
        if (!hip_ok(hipMalloc(&g_mm_st, kMsMulti * sizeof(SolverState)), "multi states") ||
            !hip_ok(hipHostMalloc(&g_mm_st_host, kMsMulti * sizeof(SolverState), hipHostMallocDefault),
                    "multi states") ||
            !hip_ok(hipMalloc(&g_mm_win, kMsMulti * sizeof(MsWin)), "multi windows"))
            return false;
    }
    if (g_mm_part_n < (size_t)2 * nbh) {
        if (g_mm_part) (void)hipFree(g_mm_part);
        g_mm_part = nullptr;
        g_mm_part_n = 0;
        if (!hip_ok(hipMalloc(&g_mm_part, sizeof(double) * 2 * nbh * kMsMulti), "multi partials")) return false;
        g_mm_part_n = (size_t)2 * nbh;
    }
    const size_t hb = (size_t)W * vb;
    if (g_mm_hist_bytes < (size_t)n * hb) {
        if (g_mm_hist) (void)hipFree(g_mm_hist);
        g_mm_hist = nullptr;
        g_mm_hist_bytes = 0;
        if (hipMalloc(&g_mm_hist, (size_t)n * hb) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        g_mm_hist_bytes = (size_t)n * hb;
    }
    auto ctl_k = [&](int k, int pred, int iter = 0) {
        SolverCtl c;
        c.st = g_mm_st + k;
        c.partials = g_mm_part + (size_t)k * g_mm_part_n;
        c.pred = pred;
        c.iter = iter;
        return c;
    };
    auto hist = [&](int k) { return (char*)g_mm_hist + (size_t)k * hb; };
    auto slot = [&](int k, int j) { return (double2*)(hist(k) + (size_t)((j - 1) % W) * vb); };

    // x_k = 0 ; r = p = b ; p_k = b ; states as multishift_solve sets them
    double sn[kMsMulti];
    for (int k = 0; k < n; k++) {
        const MsSys& d = sy[k];
        for (int i = 0; i < d.order; i++) {
            if (!hip_ok(hipMemsetAsync((char*)d.out + i * vb, 0, vb, S()), "zero x_k") ||
                !hip_ok(hipMemcpyAsync((char*)d.shiftferm + i * vb, d.b, vb, hipMemcpyDeviceToDevice, S()),
                        "p_k = b"))
                return false;
        }
        if (!hip_ok(hipMemcpyAsync(d.r, d.b, vb, hipMemcpyDeviceToDevice, S()), "r = b") ||
            !hip_ok(hipMemcpyAsync(d.p, d.b, vb, hipMemcpyDeviceToDevice, S()), "p = b") ||
            !reduce_to_host(0, d.b, nullptr, 0, &sn[k], nullptr))
            return false;
        SolverState& s = g_mm_st_host[k];
        std::memset(&s, 0, sizeof s);
        s.res = d.residuo;
        s.safety = 0.95;
        s.max_cg = d.max_cg;
        s.restart_every = 0;
        s.status = kRunning;
        s.delta = sn[k];
        s.source_norm = sn[k];
        s.omega = 1.0;
        s.gammag = 0.0;
        s.order = d.order;
        s.flag_cur = 0;
        for (int i = 0; i < d.order; i++) {
            s.flags[1][i] = 1;
            s.zeta_i[i] = 1.0;
            s.zeta_ii[i] = 1.0;
            s.gammas[i] = 0.0;
            s.shifts[i] = d.shifts[i];
        }
        s.maxiter_next = d.order;
        if (d.order == 0) s.status = kDone;
    }
    if (!hip_ok(hipMemcpyAsync(g_mm_st, g_mm_st_host, n * sizeof(SolverState), hipMemcpyHostToDevice, S()),
                "multi states") ||
        !hip_ok(hipStreamSynchronize(S()), "multi states"))
        return false;
    for (int k = 0; k < n; k++)
        hipLaunchKernelGGL(k_msw_reset, dim3(1), dim3(kBlock), 0, S(), g_mm_win + k, ctl_k(k, kAlways));

    MsMultiHop a;
    for (int l = 0; l < 8; l++) {
        a.u[l] = (const double2*)((const char*)u + l * L.su3_bytes());
        a.ph[l] = (const double*)((const char*)ph + l * L.dsoa_bytes());
    }
    a.g.nd0 = L.nd[0];
    a.g.nd0h = L.nd[0] / 2;
    a.g.nd1 = L.nd[1];
    a.g.nd2 = L.nd[2];
    a.g.nd3 = L.nd[3];
    a.g.sizeh = (int)sizeh;
    a.tmajor = (sizeh / L.nd[3]) % kBlock == 0 ? 1 : 0;  // as k_dslash<double>
    a.n = n;
    a.pred = kIfRunning;
    for (int k = 0; k < kMsMulti; k++) {
        const int q = k < n ? k : 0;
        a.st[k] = g_mm_st + q;
        a.part[k] = g_mm_part + (size_t)q * g_mm_part_n;
        a.m2[k] = sy[q].mass * sy[q].mass;
    }
    auto iteration = [&](int j) -> bool {
        // h_k = D_oe p_k ; s_k = m2_k p_k - D_eo h_k, <p_k, s_k>
        for (int k = 0; k < kMsMulti; k++) {
            const int q = k < n ? k : 0;
            a.in[k] = (const double2*)sy[q].p;
            a.out[k] = (double2*)sy[q].h;
            a.x[k] = nullptr;
        }
        a.rev = ctx().dslash_serpentine ? (int)(ctx().dslash_launches++ & 1u) : 0;
        launch_multi_hop<1>(a, nbh);
        for (int k = 0; k < kMsMulti; k++) {
            const int q = k < n ? k : 0;
            a.in[k] = (const double2*)sy[q].h;
            a.out[k] = (double2*)sy[q].s;
            a.x[k] = (const double2*)sy[q].p;
        }
        a.rev = ctx().dslash_serpentine ? (int)(ctx().dslash_launches++ & 1u) : 0;
        launch_multi_hop<0>(a, nbh);
        for (int k = 0; k < n; k++) {
            if (!launch_finalize(kFinMsAlpha, ctl_k(k, kIfRunning), nbh)) return false;
            const double2* r_old = j == 1 ? (const double2*)sy[k].r : slot(k, j - 1);
            hipLaunchKernelGGL((k_msw_r<double2>), dim3(nbh), dim3(kBlock), 0, S(), r_old, slot(k, j),
                               (const double2*)sy[k].s, (int)sizeh, ctl_k(k, kIfRunning));
            if (!launch_finalize(kFinMsLambda, ctl_k(k, kIfRunning), nbh)) return false;
            hipLaunchKernelGGL((k_msw_p<double2>), dim3(nb3), dim3(kBlock), 0, S(), (double2*)sy[k].p,
                               (const double2*)slot(k, j), n3, g_mm_win + k, ctl_k(k, kIfIter, j),
                               (double2*)nullptr);
            if (j % W == 0) {
                hipLaunchKernelGGL((k_msw_flush<double2, double2>), dim3(nb3), dim3(kBlock), 0, S(),
                                   (double2*)sy[k].out, (double2*)sy[k].shiftferm, (const double2*)hist(k), n3,
                                   sy[k].order, (const MsWin*)(g_mm_win + k), ctl_k(k, kIfIter, j));
                hipLaunchKernelGGL(k_msw_reset, dim3(1), dim3(kBlock), 0, S(), g_mm_win + k, ctl_k(k, kIfIter, j));
            }
        }
        return check_launch("multi multishift iteration");
    };
    // batches as run_batches, sized by the slowest system
    int batch = 4, last_total = -1;
    double prev_d = NAN;
    int last_cg = -1;
    for (;;) {
        if (!hip_ok(hipMemcpyAsync(g_mm_st_host, g_mm_st, n * sizeof(SolverState), hipMemcpyDeviceToHost, S()),
                    "multi states") ||
            !hip_ok(hipStreamSynchronize(S()), "multi states"))
            return false;
        int running = 0, total = 0, cg0 = 0;
        double d = -INFINITY;
        for (int k = 0; k < n; k++) {
            const SolverState& s = g_mm_st_host[k];
            total += s.cg;
            if (s.status == kRunning) {
                running++;
                cg0 = std::max(cg0, s.cg);
                const double dk = dist_ms(s);
                d = std::isfinite(dk) ? std::max(d, dk) : d;
            }
        }
        if (!running) break;
        if (total == last_total) return fail("multi multishift: solver states did not advance (iteration %d)", cg0);
        last_total = total;
        int nq = batch;
        if (std::isfinite(d) && std::isfinite(prev_d) && last_cg >= 0 && cg0 > last_cg) {
            const double rate = (prev_d - d) / (double)(cg0 - last_cg);
            if (rate > 1e-9) {
                const double left = d > 0.0 ? d / rate : 0.0;
                nq = std::min(nq, std::max(2, (int)std::ceil(1.1 * left) + 1));
            }
        }
        prev_d = std::isfinite(d) ? d : NAN;
        last_cg = cg0;
        for (int j = 1; j <= nq; j++)
            if (!iteration(cg0 + j)) return false;
        batch = std::min(2 * batch, 64);
    }
    // the partial last windows, final residuals, the per-shift check
    for (int k = 0; k < n; k++) {
        MsSys& d = sy[k];
        const int cg = g_mm_st_host[k].cg;
        d.cg = cg;
        hipLaunchKernelGGL((k_msw_flush<double2, double2>), dim3(nb3), dim3(kBlock), 0, S(), (double2*)d.out,
                           (double2*)d.shiftferm, (const double2*)hist(k), n3, d.order,
                           (const MsWin*)(g_mm_win + k), ctl_k(k, kAlways));
        if (cg > 0 && !hip_ok(hipMemcpyAsync(d.r, slot(k, cg), vb, hipMemcpyDeviceToDevice, S()), "final residual"))
            return false;
        if (!check_launch("multi multishift window flush")) return false;
        multishift_invert_iterations += cg;
        if (cg == d.max_cg) printf("WARNING: maximum number of iterations reached in invert\n");
        int check = 1;
        if (d.order > 0) {
            double g[kMaxShifts];
            if (!verify_shifts_f64(u, ph, d.b, d.out, d.order, d.mass * d.mass, d.shifts, sn[k], g)) return false;
            for (int i = 0; i < d.order; i++) check *= (g[i] <= 1) ? 1 : 0;
        }
        if (!hip_ok(hipMemcpyAsync(d.p, (const char*)d.out + (d.order > 0 ? (d.order - 1) : 0) * vb, vb,
                                   hipMemcpyDeviceToDevice, S()),
                    "loc_p"))
            return false;
        if (verbosity_lv > 0)
            printf("Terminated multishift_invert ( target res = %1.1e,source_norm = %1.1e )\tCG count %d\n",
                   d.residuo, sn[k], cg);
        d.status = check == 1 ? INVERTER_SUCCESS : INVERTER_FAILURE;
    }
    return hip_ok(hipStreamSynchronize(S()), "synchronize");
}

bool recombine(const void* xs, const void* in, void* out, const RationalApprox* ra) {
    RecombineCoeffs c;
    c.a0 = ra->RA_a0;
    c.order = ra->approx_order;
    if (c.order < 0 || c.order > kMaxShifts) return fail("recombine: order %d out of range", c.order);
    for (int k = 0; k < kMaxShifts; k++) c.a[k] = k < c.order ? ra->RA_a[k] : 0.0;
    const long n3 = 3 * ctx().lat.sizeh;
    hipLaunchKernelGGL(k_recombine, dim3(blocks_for(n3)), dim3(kBlock), 0, S(), (const double2*)xs,
                       (const double2*)in, (double2*)out, n3, n3, c);
    return check_launch("k_recombine") && hip_ok(hipStreamSynchronize(S()), "synchronize");
}

// library-owned FP32 scratch for the single-precision-accelerated multishift
// (the reference's globals aux1_f and ferm_shiftmulti_acc_f, alloc_vars.c)
static void* g_aux1_f = nullptr;
static void* g_shiftmulti_f = nullptr;
static size_t g_aux1_bytes = 0, g_shiftmulti_bytes = 0;

// inverter_multishift_wrapper's reduced-precision mode (ostp_set_multishift_precision)
// and the FP32 copies it makes when the caller's package carries none
static int g_ms_precision = 0;
static int g_ms_force_only = 0;  // the mode applies inside fermion_force_soloopenacc only
static void *g_wr_conf_f = nullptr, *g_wr_ph_f = nullptr, *g_wr_vec_f = nullptr;
static size_t g_wr_conf_f_bytes = 0, g_wr_ph_f_bytes = 0, g_wr_vec_f_bytes = 0;

static void* grow_buffer(void*& p, size_t& have, size_t need) {
    if (have >= need) return p;
    if (p) (void)hipFree(p);
    p = nullptr;
    have = 0;
    if (!hip_ok(hipMalloc(&p, need), "scratch allocation")) return nullptr;
    have = need;
    return p;
}

// ------------------------------------------------ allocation-free CG scratch
// ker_invert_openacc_noalloc keeps the reduction partials and the solver state
// of its kernels in this file-scope device array (one copy per GPU, held
// until the process ends) instead of the buffers ostp_init allocates.  Host
// code must not pass `scratch` itself to a kernel or a copy: its address on
// each device comes from hipGetSymbolAddress (noalloc_setup, called by
// ostp_init with the device current) and is kept in g_scratch_dev.
constexpr long kNoAllocMaxSizeh = 1L << 23;  // half-sites: a 64^4 lattice
constexpr size_t kNoAllocPartialBytes = sizeof(double) * 4 * (size_t)((kNoAllocMaxSizeh + 63) / 64 + 64);
constexpr size_t kNoAllocBytes = kNoAllocPartialBytes + ((sizeof(SolverState) + 255) / 256) * 256;
static_assert(kNoAllocPartialBytes % 256 == 0, "state offset alignment");
static __device__ __align__(256) char scratch[kNoAllocBytes];
constexpr int kNoAllocMaxDevices = 64;
static void* g_scratch_dev[kNoAllocMaxDevices] = {};

bool noalloc_setup(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return fail("device %d beyond %d", device, kNoAllocMaxDevices);
    if (g_scratch_dev[device]) return true;
    void* p = nullptr;
    if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return false;
    g_scratch_dev[device] = p;
    return true;
}

// For the lifetime of the object, the solver context uses the static scratch
// of the current device as its partials and state buffers.
struct ScratchSwap {
    double* partials;
    SolverState* state;
    int max_blocks;
    explicit ScratchSwap(void* base) : partials(ctx().partials), state(ctx().state), max_blocks(ctx().max_blocks) {
        ctx().partials = (double*)base;
        ctx().state = (SolverState*)((char*)base + kNoAllocPartialBytes);
        ctx().max_blocks = (int)((kNoAllocMaxSizeh + 63) / 64 + 64);
    }
    ~ScratchSwap() {
        ctx().partials = partials;
        ctx().state = state;
        ctx().max_blocks = max_blocks;
    }
};

// power_method for n <= kMsMulti start vectors p_k on the operators
// coef_k - Deo Doe that differ only by coef_k (the flavours of one gauge
// field): the two hops of each iteration run once for all of them
// (k_ms_multi_hop), the scaling and the norm per system on its own state, so
// each system's iterations and norms are those of power_method.
static bool power_method_multi(const void* u, const void* ph, const double* coef, void* const* r, void* const* h,
                               void* const* p, int n, double* lambda, int* iters) {
    const Lattice& L = ctx().lat;
    const long sizeh = L.sizeh, n3 = 3 * sizeh;
    const int nbh = blocks_for(sizeh);
    if (n < 1 || n > kMsMulti) return false;
    if (!g_mm_st) {
        if (!hip_ok(hipMalloc(&g_mm_st, kMsMulti * sizeof(SolverState)), "multi states") ||
            !hip_ok(hipHostMalloc(&g_mm_st_host, kMsMulti * sizeof(SolverState), hipHostMallocDefault),
                    "multi states") ||
            !hip_ok(hipMalloc(&g_mm_win, kMsMulti * sizeof(MsWin)), "multi windows"))
            return false;
    }
    if (g_mm_part_n < (size_t)2 * nbh) {
        if (g_mm_part) (void)hipFree(g_mm_part);
        g_mm_part = nullptr;
        g_mm_part_n = 0;
        if (!hip_ok(hipMalloc(&g_mm_part, sizeof(double) * 2 * nbh * kMsMulti), "multi partials")) return false;
        g_mm_part_n = (size_t)2 * nbh;
    }
    auto ctl_k = [&](int k, int pred) {
        SolverCtl c;
        c.st = g_mm_st + k;
        c.partials = g_mm_part + (size_t)k * g_mm_part_n;
        c.pred = pred;
        return c;
    };
    for (int k = 0; k < n; k++) {
        double sn = 0.0;
        if (!reduce_to_host(0, p[k], nullptr, 0, &sn, nullptr)) return false;
        SolverState& s = g_mm_st_host[k];
        std::memset(&s, 0, sizeof s);
        s.max_cg = 1000000;  // the reference has no iteration limit
        s.status = kRunning;
        s.lambda = std::sqrt(sn);
    }
    if (!hip_ok(hipMemcpyAsync(g_mm_st, g_mm_st_host, n * sizeof(SolverState), hipMemcpyHostToDevice, S()),
                "multi states") ||
        !hip_ok(hipStreamSynchronize(S()), "multi states"))
        return false;
    MsMultiHop a;
    for (int l = 0; l < 8; l++) {
        a.u[l] = (const double2*)((const char*)u + l * L.su3_bytes());
        a.ph[l] = (const double*)((const char*)ph + l * L.dsoa_bytes());
    }
    a.g.nd0 = L.nd[0];
    a.g.nd0h = L.nd[0] / 2;
    a.g.nd1 = L.nd[1];
    a.g.nd2 = L.nd[2];
    a.g.nd3 = L.nd[3];
    a.g.sizeh = (int)sizeh;
    a.tmajor = (sizeh / L.nd[3]) % kBlock == 0 ? 1 : 0;
    a.n = n;
    a.pred = kIfRunning;
    for (int k = 0; k < kMsMulti; k++) {
        const int q = k < n ? k : 0;
        a.st[k] = g_mm_st + q;
        a.part[k] = g_mm_part + (size_t)q * g_mm_part_n;
        a.m2[k] = coef[q];
    }
    auto iteration = [&]() -> bool {
        // r = p / norm ; p = coef r - Deo Doe r ; norm = |p|, stop test
        for (int k = 0; k < n; k++)
            hipLaunchKernelGGL(k_pm_scale, dim3(blocks_for(n3)), dim3(kBlock), 0, S(), (const double2*)p[k],
                               (double2*)r[k], n3, ctl_k(k, kIfRunning));
        for (int k = 0; k < kMsMulti; k++) {
            const int q = k < n ? k : 0;
            a.in[k] = (const double2*)r[q];
            a.out[k] = (double2*)h[q];
            a.x[k] = nullptr;
        }
        a.rev = ctx().dslash_serpentine ? (int)(ctx().dslash_launches++ & 1u) : 0;
        launch_multi_hop<1>(a, nbh);
        for (int k = 0; k < kMsMulti; k++) {
            const int q = k < n ? k : 0;
            a.in[k] = (const double2*)h[q];
            a.out[k] = (double2*)p[q];
            a.x[k] = (const double2*)r[q];
        }
        a.rev = ctx().dslash_serpentine ? (int)(ctx().dslash_launches++ & 1u) : 0;
        launch_multi_hop<0>(a, nbh);
        for (int k = 0; k < n; k++)
            if (!launch_norm2<double2, kFinPower>(p[k], ctl_k(k, kIfRunning))) return false;
        return check_launch("multi power method");
    };
    int batch = 4, last_total = -1;
    for (;;) {
        if (!hip_ok(hipMemcpyAsync(g_mm_st_host, g_mm_st, n * sizeof(SolverState), hipMemcpyDeviceToHost, S()),
                    "multi states") ||
            !hip_ok(hipStreamSynchronize(S()), "multi states"))
            return false;
        int running = 0, total = 0;
        for (int k = 0; k < n; k++) {
            total += g_mm_st_host[k].cg;
            if (g_mm_st_host[k].status == kRunning) running++;
        }
        if (!running) break;
        if (total == last_total) return fail("multi power method: states did not advance");
        last_total = total;
        for (int j = 0; j < batch; j++)
            if (!iteration()) return false;
        batch = std::min(2 * batch, 64);
    }
    for (int k = 0; k < n; k++) {
        lambda[k] = g_mm_st_host[k].lambda;
        if (iters) iters[k] = g_mm_st_host[k].cg;
    }
    return true;
}

// 1 where two FP64 arrays of n doubles differ in any bit
__global__ __launch_bounds__(kBlock) void k_differ(const unsigned long long* __restrict__ a,
                                                   const unsigned long long* __restrict__ b, long n, int* flag) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n && a[i] != b[i]) *flag = 1;
}
static int* g_differ_flag = nullptr;

// the U(1) phases of two flavours are the same (one array, or equal values)
static bool same_phases(const void* a, const void* b) {
    if (a == b) return true;
    const long n = 8 * ctx().lat.sizeh;
    if (!g_differ_flag && !hip_ok(hipMalloc(&g_differ_flag, sizeof(int)), "phase comparison")) return false;
    int h = 0;
    if (!hip_ok(hipMemsetAsync(g_differ_flag, 0, sizeof(int), S()), "phase comparison")) return false;
    hipLaunchKernelGGL(k_differ, dim3(blocks_for(n)), dim3(kBlock), 0, S(), (const unsigned long long*)a,
                       (const unsigned long long*)b, n, g_differ_flag);
    if (!check_launch("k_differ") ||
        !hip_ok(hipMemcpyAsync(&h, g_differ_flag, sizeof(int), hipMemcpyDeviceToHost, S()), "phase comparison") ||
        !hip_ok(hipStreamSynchronize(S()), "phase comparison"))
        return false;
    return h == 0;
}

static void* g_many_work = nullptr;
static size_t g_many_work_bytes = 0;
static int g_ms_many = 1;  // rational solves of several sources together where they share the operator

// The rational solves of n sources (in_k with the approximation ras[k] and
// the mass of pars[k], solutions to out_k, order_k vectors each) as
// inverter_multishift_wrapper would run them one by one in FP64, run
// together (multishift_solve_multi, in groups of kMsMulti) when every source
// has the same U(1) phases.  false, with nothing solved, when that does not
// apply (another precision mode, the single-precision accelerated path,
// fewer than two sources, different phases, or no memory for it).
bool multishift_fp64_many(const void* u, ferm_param* const* pars, const RationalApprox* const* ras,
                          const void* const* in, void* const* out, int n, double res, int max_cg,
                          int convergence_importance) {
    if (!g_ms_many || n < 2 || g_ms_precision != 0 || inverter_tricks.singlePInvAccelMultiInv) return false;
    const unsigned long e0 = error_count();
    const Lattice& L = ctx().lat;
    const size_t vb = L.vec3_bytes();
    const char* nm = "inverter_multishift_wrapper";
    const void* ph0 = resolve(pars[0]->phases, 8 * L.dsoa_bytes(), nm);
    if (!ph0) return false;
    int maxo = 0;
    for (int k = 0; k < n; k++) {
        const void* ph = resolve(pars[k]->phases, 8 * L.dsoa_bytes(), nm);
        const int order = ras[k]->approx_order;
        if (!ph || order < 0 || order > kMaxShifts || !same_phases(ph0, ph)) return false;
        maxo = std::max(maxo, order);
    }
    // per system of a group: r, h, s, p and the search directions p_k
    const size_t per = (size_t)(4 + std::max(maxo, 1)) * vb;
    char* w = (char*)grow_buffer(g_many_work, g_many_work_bytes, kMsMulti * per);
    if (!w) {
        (void)hipGetLastError();
        forgive_errors(e0);
        return false;
    }
    for (int k0 = 0; k0 < n; k0 += kMsMulti) {
        const int m = std::min(kMsMulti, n - k0);
        MsSys sy[kMsMulti];
        for (int r = 0; r < m; r++) {
            ferm_param* fp = pars[k0 + r];
            char* base = w + (size_t)r * per;
            MsSys& d = sy[r];
            d.b = in[k0 + r];
            d.out = out[k0 + r];
            d.r = base;
            d.h = base + vb;
            d.s = base + 2 * vb;
            d.p = base + 3 * vb;
            d.shiftferm = base + 4 * vb;
            d.mass = fp->ferm_mass;
            d.order = ras[k0 + r]->approx_order;
            d.shifts = ras[k0 + r]->RA_b;
            d.residuo = res;
            d.max_cg = max_cg;
        }
        if (!multishift_solve_multi(u, ph0, sy, m)) {
            if (error_count() != e0) return false;
            // no window memory for the group: one by one, as the wrapper does
            for (int r = 0; r < m; r++) {
                MsBufs B;
                B.u = u;
                B.ph = ph0;
                B.b = sy[r].b;
                B.out = sy[r].out;
                B.shiftferm = sy[r].shiftferm;
                B.r = sy[r].r;
                B.h = sy[r].h;
                B.s = sy[r].s;
                B.p = sy[r].p;
                int cg = 0;
                sy[r].status = multishift_solve<double>(B, sy[r].mass, sy[r].order, sy[r].shifts, res, max_cg, &cg,
                                                        nullptr);
                sy[r].cg = cg;
            }
        }
        for (int r = 0; r < m; r++) convergence_messages(convergence_importance, sy[r].status);
    }
    return error_count() == e0;
}

// ker_find_max_eigenvalue_openacc for n flavours (start vectors p_k, work
// r_k, h_k) together when their U(1) phases agree; false, with nothing done,
// otherwise (or when batching is off), so the caller runs them one by one.
bool find_max_eigenvalues_many(const void* u, ferm_param* const* pars, void* const* r, void* const* h,
                               void* const* p, int n, double* maxv) {
    if (!g_ms_many || n < 2 || n > kMsMulti) return false;
    const Lattice& L = ctx().lat;
    const void* ph0 = resolve(pars[0]->phases, 8 * L.dsoa_bytes(), "find_min_max_eigenvalue_soloopenacc");
    if (!ph0) return false;
    for (int k = 1; k < n; k++) {
        const void* ph = resolve(pars[k]->phases, 8 * L.dsoa_bytes(), "find_min_max_eigenvalue_soloopenacc");
        if (!ph || !same_phases(ph0, ph)) return false;
    }
    double coef[kMsMulti];
    for (int k = 0; k < n; k++) coef[k] = pars[k]->ferm_mass * pars[k]->ferm_mass;
    return power_method_multi(u, ph0, coef, r, h, p, n, maxv, nullptr);
}

}  // namespace ostp

using namespace ostp;

// ------------------------------------------------------------ C entry points
extern "C" {

int ker_invert_openacc(const su3_soa* u, ferm_param* pars, vec3_soa* solution, const vec3_soa* in, double res,
                       vec3_soa* loc_r, vec3_soa* loc_h, vec3_soa* loc_s, vec3_soa* loc_p, const int max_cg,
                       double shift, int* cg_return) {
    const char* nm = "ker_invert_openacc";
    if (!require_init(nm)) return INVERTER_FAILURE;
    if (!pars || !cg_return) return fail("%s: null argument", nm) ? 0 : INVERTER_FAILURE;
    const Lattice& L = ctx().lat;
    const size_t vb = L.vec3_bytes();
    CgBufs B;
    B.u = resolve(u, 8 * L.su3_bytes(), nm);
    B.ph = resolve(pars->phases, 8 * L.dsoa_bytes(), nm);
    B.b = resolve(in, vb, nm);
    B.x = resolve(solution, vb, nm);
    B.r = resolve(loc_r, vb, nm);
    B.h = resolve(loc_h, vb, nm);
    B.s = resolve(loc_s, vb, nm);
    B.p = resolve(loc_p, vb, nm);
    if (!B.u || !B.ph || !B.b || !B.x || !B.r || !B.h || !B.s || !B.p) return INVERTER_FAILURE;
    int cg = 0;
    int st = cg_solve(B, pars->ferm_mass, shift, res, max_cg, inverter_tricks.restartingEvery, &cg, nullptr);
    *cg_return = cg;
    return st;
}

int ker_invert_openacc_noalloc(const su3_soa* u, ferm_param* pars, vec3_soa* solution, const vec3_soa* in,
                               double res, vec3_soa* loc_r, vec3_soa* loc_h, vec3_soa* loc_s, vec3_soa* loc_p,
                               const int max_cg, double shift, int* cg_return) {
    const char* nm = "ker_invert_openacc_noalloc";
    if (!require_init(nm)) return INVERTER_FAILURE;
    if (ctx().lat.sizeh > kNoAllocMaxSizeh) {
        fail("%s: the lattice has %ld half-sites; the static scratch holds at most %ld (a 64^4 lattice)", nm,
             ctx().lat.sizeh, kNoAllocMaxSizeh);
        return INVERTER_FAILURE;
    }
    void* base = g_scratch_dev[ctx().device];
    if (!base) {
        fail("%s: no scratch address for device %d (ostp_init sets it up)", nm, ctx().device);
        return INVERTER_FAILURE;
    }
    ScratchSwap swap(base);
    return ker_invert_openacc(u, pars, solution, in, res, loc_r, loc_h, loc_s, loc_p, max_cg, shift, cg_return);
}

int inverter_mixed_precision(inverter_package ip, ferm_param* pars, vec3_soa* solution, const vec3_soa* in,
                             double res, const int max_cg, double shift, int* cg_return) {
    const char* nm = "inverter_mixed_precision";
    if (!require_init(nm)) return INVERTER_FAILURE;
    if (!pars || !cg_return) return fail("%s: null argument", nm) ? 0 : INVERTER_FAILURE;
    const Lattice& L = ctx().lat;
    const size_t vb = L.vec3_bytes(), vbf = L.vec3f_bytes();
    MxBufs B;
    B.u = resolve(ip.u, 8 * L.su3_bytes(), nm);
    B.u_f = resolve(ip.u_f, 8 * L.su3f_bytes(), nm);
    B.ph = resolve(pars->phases, 8 * L.dsoa_bytes(), nm);
    B.ph_f = resolve(pars->phases_f, 8 * L.fsoa_bytes(), nm);
    B.b = resolve(in, vb, nm);
    B.x = resolve(solution, vb, nm);
    B.r = resolve(ip.loc_r, vb, nm);
    B.h = resolve(ip.loc_h, vb, nm);
    B.s = resolve(ip.loc_s, vb, nm);
    B.rf = resolve(ip.loc_r_f, vbf, nm);
    B.hf = resolve(ip.loc_h_f, vbf, nm);
    B.sf = resolve(ip.loc_s_f, vbf, nm);
    B.pf = resolve(ip.loc_p_f, vbf, nm);
    B.outf = resolve(ip.out_f, vbf, nm);
    if (!B.u || !B.u_f || !B.ph || !B.ph_f || !B.b || !B.x || !B.r || !B.h || !B.s || !B.rf || !B.hf || !B.sf ||
        !B.pf || !B.outf)
        return INVERTER_FAILURE;
    int cg = 0;
    int st = mixed_solve(B, pars->ferm_mass, shift, res, max_cg, inverter_tricks.mixedPrecisionDelta, &cg, nullptr,
                         nullptr);
    *cg_return = cg;
    return st;
}

int multishift_invert(const su3_soa* u, ferm_param* pars, RationalApprox* approx, vec3_soa* out, const vec3_soa* in,
                      double residuo, vec3_soa* loc_r, vec3_soa* loc_h, vec3_soa* loc_s, vec3_soa* loc_p,
                      vec3_soa* shiftferm, const int max_cg, int* cg_return) {
    const char* nm = "multishift_invert";
    if (!require_init(nm)) return INVERTER_FAILURE;
    if (!pars || !approx || !cg_return) return fail("%s: null argument", nm) ? 0 : INVERTER_FAILURE;
    const int order = approx->approx_order;
    if (order < 0 || order > kMaxShifts) return fail("%s: approx_order %d", nm, order) ? 0 : INVERTER_FAILURE;
    const Lattice& L = ctx().lat;
    const size_t vb = L.vec3_bytes();
    const size_t nv = order > 0 ? (size_t)order : 1;
    MsBufs B;
    B.u = resolve(u, 8 * L.su3_bytes(), nm);
    B.ph = resolve(pars->phases, 8 * L.dsoa_bytes(), nm);
    B.b = resolve(in, vb, nm);
    B.out = resolve(out, nv * vb, nm);
    B.shiftferm = resolve(shiftferm, nv * vb, nm);
    B.r = resolve(loc_r, vb, nm);
    B.h = resolve(loc_h, vb, nm);
    B.s = resolve(loc_s, vb, nm);
    B.p = resolve(loc_p, vb, nm);
    if (!B.u || !B.ph || !B.b || !B.out || !B.shiftferm || !B.r || !B.h || !B.s || !B.p) return INVERTER_FAILURE;
    int cg = 0;
    int st = multishift_solve<double>(B, pars->ferm_mass, order, approx->RA_b, residuo, max_cg, &cg, nullptr);
    *cg_return = cg;
    return st;
}

void recombine_shifted_vec3_to_vec3(const vec3_soa* in_shifted, const vec3_soa* in, vec3_soa* out,
                                    const RationalApprox* approx) {
    const char* nm = "recombine_shifted_vec3_to_vec3";
    if (!require_init(nm)) return;
    if (!approx) { fail("%s: null approx", nm); return; }
    const Lattice& L = ctx().lat;
    const size_t vb = L.vec3_bytes();
    const size_t nv = approx->approx_order > 0 ? (size_t)approx->approx_order : 1;
    const void* xs = resolve(in_shifted, nv * vb, nm);
    const void* di = resolve(in, vb, nm);
    void* dout = resolve(out, vb, nm);
    if (xs && di && dout) recombine(xs, di, dout, approx);
}

void convergence_messages(int conv_importance, int inverter_status) {
    if (INVERTER_FAILURE == inverter_status) {
        if (CONVERGENCE_CRITICAL == conv_importance) {
            printf("\n\n\t\tERROR : inverter failed to converge in a critical region of the code. Exiting now!!\n\n");
            exit(1);
        }
        printf("\n\t\tWARNING : inverter failed to converge.\n");
    }
}

int inverter_wrapper(inverter_package ip, ferm_param* pars, vec3_soa* out, const vec3_soa* in, double res,
                     int max_cg, double shift, int convergence_importance) {
    int cg_return = 0;
    int st;
    if (inverter_tricks.useMixedPrecision)
        st = inverter_mixed_precision(ip, pars, out, in, res, max_cg, shift, &cg_return);
    else
        st = ker_invert_openacc(ip.u, pars, out, in, res, ip.loc_r, ip.loc_h, ip.loc_s, ip.loc_p, max_cg, shift,
                                &cg_return);
    convergence_messages(convergence_importance, st);
    return cg_return;
}

// Meas/ferm_meas.c:50-71
void eo_inversion(inverter_package ip, ferm_param* tfermions_parameters, double res, int max_cg, vec3_soa* in_e,
                  vec3_soa* in_o, vec3_soa* out_e, vec3_soa* out_o, vec3_soa* phi_e, vec3_soa* phi_o) {
    if (!require_init("eo_inversion")) return;
    if (!tfermions_parameters) {
        fail("eo_inversion: null ferm_param");
        return;
    }
    const double m = tfermions_parameters->ferm_mass;
    acc_Deo(ip.u, phi_e, in_o, tfermions_parameters->phases);
    combine_in1_x_fact1_minus_in2_back_into_in2(in_e, m, phi_e);
    set_vec3_soa_to_zero(out_e);
    inverter_wrapper(ip, tfermions_parameters, out_e, phi_e, res, max_cg, 0, CONVERGENCE_CRITICAL);
    acc_Doe(ip.u, phi_o, out_e, tfermions_parameters->phases);
    combine_in1_minus_in2_allxfact(in_o, phi_o, (double)1 / m, out_o);
}

// Z2-noise chiral condensate (Meas/ferm_meas.c:282-320) for nvec noise vectors
// at once: eo_inversion's even-odd split for every vector, the even systems
// (m^2 - Deo Doe) x_e = m eta_e - Deo eta_o solved together by the batched
// solver, x_o = (eta_o - Doe x_e) / m, and per vector
// (<eta_e, x_e> + <eta_o, x_o>) * degeneracy * 0.25 / V.
int ostp_chiral_condensate(const su3_soa* u, const su3_soa_f* u_f, ferm_param* pars, int nvec, uint64_t seed,
                           double degeneracy, double res, int max_cg, int half_inner, double* cond,
                           int* cg_return) {
    const char* nm = "ostp_chiral_condensate";
    if (!require_init(nm)) return 0;
    if (nvec < 1 || !pars || !cond) {
        fail("%s: invalid arguments", nm);
        return 0;
    }
    const Lattice& L = ctx().lat;
    const size_t vb = L.vec3_bytes();
    // per vector: eta_e, eta_o, phi (source, then temporary), x_e, x_o
    char* buf = nullptr;
    if (!hip_ok(hipMalloc(&buf, 5 * vb * (size_t)nvec), "chiral condensate vectors")) return 0;
    auto vec = [&](int k, int j) { return (vec3_soa*)(buf + ((size_t)k * 5 + j) * vb); };
    const double m = pars->ferm_mass;
    std::vector<ostp_batch_system> sys(nvec);
    for (int k = 0; k < nvec; k++) {
        vec3_soa *ee = vec(k, 0), *eo = vec(k, 1), *phi = vec(k, 2), *xe = vec(k, 3);
        ostp_generate_vec3_soa_z2noise(ee, seed, 2u * (uint32_t)k);
        ostp_generate_vec3_soa_z2noise(eo, seed, 2u * (uint32_t)k + 1u);
        acc_Deo(u, phi, eo, pars->phases);
        combine_in1_x_fact1_minus_in2_back_into_in2(ee, m, phi);
        set_vec3_soa_to_zero(xe);
        ostp_batch_system& q = sys[k];
        for (int d = 0; d < 4; d++) q.nd[d] = L.nd[d];
        q.par_temp = L.par_temp;
        q.u = u;
        q.u_f = u_f;
        q.phases = pars->phases;
        q.phases_f = pars->phases_f;
        q.in = phi;
        q.out = xe;
        q.mass = m;
    }
    const int conv = batched_solve_internal(nvec, sys.data(), res, max_cg, 0.1, g_internal_ipg, cg_return, nullptr, nullptr,
                                            half_inner != 0);
    const double factor = degeneracy * 0.25 / ((double)L.nd[0] * L.nd[1] * L.nd[2] * L.nd[3]);
    bool ok = true;
    for (int k = 0; ok && k < nvec; k++) {
        vec3_soa *ee = vec(k, 0), *eo = vec(k, 1), *tmp = vec(k, 2), *xe = vec(k, 3), *xo = vec(k, 4);
        acc_Doe(u, tmp, xe, pars->phases);
        combine_in1_minus_in2_allxfact(eo, tmp, (double)1 / m, xo);
        double re_e = 0, im_e = 0, re_o = 0, im_o = 0;
        ok = reduce_to_host(2, ee, xe, 0, &re_e, &im_e) && reduce_to_host(2, eo, xo, 0, &re_o, &im_o);
        cond[2 * k] = (re_e + re_o) * factor;
        cond[2 * k + 1] = (im_e + im_o) * factor;
    }
    (void)hipFree(buf);
    return ok ? conv : 0;
}

int inverter_multishift_wrapper(inverter_package ip, ferm_param* pars, RationalApprox* approx, vec3_soa* out,
                                const vec3_soa* in, double res, int max_cg, int convergence_importance) {
    const char* nm = "inverter_multishift_wrapper";
    if (!require_init(nm)) return 0;
    int total_iterations = 0;
    int cg_return = 0;
    int st;
    if (g_ms_precision != 0 && (!g_ms_force_only || in_fermion_force())) {
        // the library's reduced-precision rational solve (ostp_set_multishift_precision),
        // with FP32 links, phases and vectors from library buffers where the package has none
        const unsigned long e0 = error_count();
        const Lattice& L = ctx().lat;
        const int order = approx->approx_order;
        const size_t vbf = L.vec3f_bytes();
        inverter_package q = ip;
        ferm_param pf = *pars;
        if (!q.u_f) {
            const void* du = resolve(ip.u, 8 * L.su3_bytes(), nm);
            void* cf = grow_buffer(g_wr_conf_f, g_wr_conf_f_bytes, 8 * L.su3f_bytes());
            if (!du || !cf) return 0;
            ostp_convert_double_to_float_gauge((const su3_soa*)du, (su3_soa_f*)cf);
            q.u_f = (const su3_soa_f*)cf;
        }
        if (!pf.phases_f) {
            void* phf = grow_buffer(g_wr_ph_f, g_wr_ph_f_bytes, 8 * L.fsoa_bytes());
            if (!phf) return 0;
            ostp_phases_to_float(pars->phases, (float_soa*)phf);
            pf.phases_f = (float_soa*)phf;
        }
        if (!q.ferm_shift_temp_f || !q.loc_r_f || !q.loc_h_f || !q.loc_s_f || !q.loc_p_f || !q.out_f) {
            char* v = (char*)grow_buffer(g_wr_vec_f, g_wr_vec_f_bytes, (size_t)(order + 5) * vbf);
            if (!v) return 0;
            q.loc_r_f = (vec3_soa_f*)v;
            q.loc_h_f = (vec3_soa_f*)(v + vbf);
            q.loc_s_f = (vec3_soa_f*)(v + 2 * vbf);
            q.loc_p_f = (vec3_soa_f*)(v + 3 * vbf);
            q.out_f = (vec3_soa_f*)(v + 4 * vbf);
            q.ferm_shift_temp_f = (vec3_soa_f*)(v + 5 * vbf);
        }
        if (error_count() != e0) return 0;
        std::vector<int> cgr(order > 0 ? order : 1, 0);
        std::vector<double> trs(order > 0 ? order : 1, 0.0);
        const int conv = ostp_multishift_mixed_solve(q, &pf, approx, out, in, res, max_cg, g_ms_precision == 2 ? 1 : 0,
                                                     &cg_return, cgr.data(), trs.data());
        convergence_messages(convergence_importance, conv == order ? INVERTER_SUCCESS : INVERTER_FAILURE);
        total_iterations = cg_return;
        for (int k = 0; k < order; k++) total_iterations += cgr[k];
        multishift_invert_iterations += cg_return;
        return total_iterations;
    }
    if (inverter_tricks.singlePInvAccelMultiInv) {
        // SP multishift to a loose target, then per-shift DP (or mixed) CG refinement
        // (inverter_wrappers.c:58-98)
        const Lattice& L = ctx().lat;
        const int order = approx->approx_order;
        const size_t nv = order > 0 ? (size_t)order : 1;
        void* aux1 = grow_buffer(g_aux1_f, g_aux1_bytes, L.vec3f_bytes());
        void* xsf = grow_buffer(g_shiftmulti_f, g_shiftmulti_bytes, nv * L.vec3f_bytes());
        const void* db = resolve(in, L.vec3_bytes(), nm);
        if (!aux1 || !xsf || !db) return 0;
        if (!launch_d2f(db, aux1, 3 * L.sizeh)) return 0;
        float target = 8e-7f * sqrtf((float)L.sizeh);
        if (target < res) target = (float)res;
        MsBufs B;
        B.u = resolve(ip.u_f, 8 * L.su3f_bytes(), nm);
        B.ph = resolve(pars->phases_f, 8 * L.fsoa_bytes(), nm);
        B.b = aux1;
        B.out = xsf;
        B.shiftferm = resolve(ip.ferm_shift_temp_f, nv * L.vec3f_bytes(), nm);
        B.r = resolve(ip.loc_r_f, L.vec3f_bytes(), nm);
        B.h = resolve(ip.loc_h_f, L.vec3f_bytes(), nm);
        B.s = resolve(ip.loc_s_f, L.vec3f_bytes(), nm);
        B.p = resolve(ip.loc_p_f, L.vec3f_bytes(), nm);
        void* dout = resolve(out, nv * L.vec3_bytes(), nm);
        if (!B.u || !B.ph || !B.shiftferm || !B.r || !B.h || !B.s || !B.p || !dout) return 0;
        st = multishift_solve<float>(B, pars->ferm_mass, order, approx->RA_b, (double)target, max_cg, &cg_return,
                                     nullptr);
        convergence_messages(convergence_importance, st);
        total_iterations += cg_return;
        for (int k = 0; k < order; k++) {
            double bshift = approx->RA_b[k];
            if (verbosity_lv > 3) printf("Shift %d, %f\n", k, bshift);
            if (!launch_f2d((const char*)xsf + k * L.vec3f_bytes(), (char*)dout + k * L.vec3_bytes(), 3 * L.sizeh))
                return 0;
            if (!hip_ok(hipStreamSynchronize(ctx().stream), nm)) return 0;
            st = inverter_wrapper(ip, pars, (vec3_soa*)((char*)out + k * L.vec3_bytes()), in, res, max_cg, bshift,
                                  convergence_importance);
            convergence_messages(convergence_importance, st);
            // the reference adds the multishift's cg_return here (inverter_wrappers.c:96)
            total_iterations += cg_return;
        }
    } else {
        st = multishift_invert(ip.u, pars, approx, out, in, res, ip.loc_r, ip.loc_h, ip.loc_s, ip.loc_p,
                               ip.ferm_shift_temp, max_cg, &cg_return);
        convergence_messages(convergence_importance, st);
        total_iterations += cg_return;
    }
    return total_iterations;
}

int ostp_set_multishift_batching(int on) {
    const int prev = g_ms_many;
    g_ms_many = on ? 1 : 0;
    return prev;
}

int ostp_set_batched_graph_iterations(int n) {
    if (n < 0 || n > 4096) {
        fail("ostp_set_batched_graph_iterations: %d outside 0..4096", n);
        return -1;
    }
    const int prev = g_internal_ipg;
    g_internal_ipg = n;
    return prev;
}

int ostp_set_multishift_precision(int mode, int force_only) {
    if (mode < 0 || mode > 2) {
        fail("ostp_set_multishift_precision: mode %d outside 0..2", mode);
        return -1;
    }
    const int prev = g_ms_precision;
    g_ms_precision = mode;
    g_ms_force_only = force_only ? 1 : 0;
    return prev;
}

// Reduced-precision rational solve.  FP32 CG-M: FP32 Dslash and vectors, the
// reference's FP64 shift recurrences (inverter_multishift_full.c), FP32
// accumulators for the shifted solutions and for the base system's solution
// x0.  Reliable updates (as in inverter_mixedp.c: the base |r|^2 has fallen
// by mixed_delta since the last one): the accumulators are added to the FP64
// solutions and the base residual is replaced by the FP64 b - M x0, so the
// Krylov space is kept and the recurrences continue (the iteration count is
// the FP64 CG-M's).  Shifted residuals are zeta_k r by construction and are
// never re-synchronised, so the FP32 search directions leave errors in the
// shifted solutions (relative residuals ~1e-8 at 32^4): every shift is then
// refined by the batched solver, all together, from its CG-M solution.
int ostp_multishift_mixed_solve(inverter_package ip, ferm_param* pars, RationalApprox* approx, vec3_soa* out,
                                const vec3_soa* in, double res, int max_cg, int half_inner, int* cg_multishift,
                                int* cg_refine, double* true_res) {
    const char* nm = "ostp_multishift_mixed_solve";
    if (!require_init(nm)) return 0;
    const int order = approx->approx_order;
    if (order < 1 || order > kMaxShifts) {
        fail("%s: approx_order %d", nm, order);
        return 0;
    }
    const Lattice& L = ctx().lat;
    const long sizeh = L.sizeh, n3 = 3 * sizeh;
    const int nbh = blocks_for(sizeh), nb3 = blocks_for(n3);
    const size_t vb = L.vec3_bytes(), vbf = L.vec3f_bytes();
    void* bf = grow_buffer(g_aux1_f, g_aux1_bytes, vbf);
    // shift updates once per window of W iterations (as multishift_solve),
    // added straight into the FP64 solutions; without the window memory, FP32
    // accumulators updated every iteration and moved into the FP64 solutions
    // at each reliable update
    const int W = msw_window(vbf);
    void* xsf = W >= 2 ? (void*)nullptr : grow_buffer(g_shiftmulti_f, g_shiftmulti_bytes, (size_t)order * vbf);
    const void* db = resolve(in, vb, nm);
    void* dout = resolve(out, (size_t)order * vb, nm);
    const void* u = resolve(ip.u, 8 * L.su3_bytes(), nm);
    const void* uf = resolve(ip.u_f, 8 * L.su3f_bytes(), nm);
    const void* ph = resolve(pars->phases, 8 * L.dsoa_bytes(), nm);
    const void* phf = resolve(pars->phases_f, 8 * L.fsoa_bytes(), nm);
    void* pk = resolve(ip.ferm_shift_temp_f, (size_t)order * vbf, nm);
    void* r = resolve(ip.loc_r_f, vbf, nm);
    void* h = resolve(ip.loc_h_f, vbf, nm);
    void* sv = resolve(ip.loc_s_f, vbf, nm);
    void* p = resolve(ip.loc_p_f, vbf, nm);
    void* x0f = resolve(ip.out_f, vbf, nm);
    void* x0 = resolve(ip.loc_p, vb, nm);
    void* h64 = resolve(ip.loc_h, vb, nm);
    void* s64 = resolve(ip.loc_s, vb, nm);
    void* r64 = resolve(ip.loc_r, vb, nm);
    if (!bf || (W < 2 && !xsf) || !db || !dout || !u || !uf || !ph || !phf || !pk || !r || !h || !sv || !p || !x0f ||
        !x0 || !h64 || !s64 || !r64)
        return 0;
    const double m2 = pars->ferm_mass * pars->ferm_mass;
    const auto t_start = std::chrono::steady_clock::now();

    // x_k = x_k_f = x0 = x0_f = 0 ; r = p = p_k = (float) b
    if (!launch_d2f(db, bf, n3)) return 0;
    bool ok = hip_ok(hipMemsetAsync(dout, 0, (size_t)order * vb, S()), "zero x_k") &&
              (!xsf || hip_ok(hipMemsetAsync(xsf, 0, (size_t)order * vbf, S()), "zero x_k_f")) &&
              hip_ok(hipMemsetAsync(x0, 0, vb, S()), "zero x0") && hip_ok(hipMemsetAsync(x0f, 0, vbf, S()), "zero x0_f") &&
              hip_ok(hipMemcpyAsync(r, bf, vbf, hipMemcpyDeviceToDevice, S()), "r = b") &&
              hip_ok(hipMemcpyAsync(p, bf, vbf, hipMemcpyDeviceToDevice, S()), "p = b");
    for (int k = 0; ok && k < order; k++)
        ok = hip_ok(hipMemcpyAsync((char*)pk + k * vbf, bf, vbf, hipMemcpyDeviceToDevice, S()), "p_k = b");
    if (!ok) return 0;
    double sn = 0.0;
    if (!reduce_to_host(0, bf, nullptr, 1, &sn, nullptr)) return 0;
    init_state(res, 0.95, max_cg, 0);
    SolverState& s0 = H();
    s0.delta = sn;
    s0.source_norm = sn;
    s0.omega = 1.0;
    s0.gammag = 0.0;
    s0.order = order;
    s0.flag_cur = 0;
    s0.mixed_delta = 0.1;
    for (int k = 0; k < order; k++) {
        s0.flags[1][k] = 1;
        s0.zeta_i[k] = 1.0;
        s0.zeta_ii[k] = 1.0;
        s0.gammas[k] = 0.0;
        s0.shifts[k] = approx->RA_b[k];
    }
    s0.maxiter_next = order;
    if (!push_state()) return 0;

    // the FP32 accumulators a reliable update moves into FP64: the base
    // system's (whose solution the FP64 residual needs), and without the
    // window the shifts' too
    MsxFlush fl;
    fl.n = xsf ? order + 1 : 1;
    fl.x64[0] = (double2*)x0;
    fl.x32[0] = (float2*)x0f;
    for (int k = 0; xsf && k < order; k++) {
        fl.x64[k + 1] = (double2*)((char*)dout + k * vb);
        fl.x32[k + 1] = (float2*)((char*)xsf + k * vbf);
    }
    float2* hist = nullptr;
    if (W >= 2) {
        if (g_msw_hist_bytes < (size_t)W * vbf) {
            if (g_msw_hist) (void)hipFree(g_msw_hist);
            g_msw_hist = nullptr;
            g_msw_hist_bytes = 0;
            if (!hip_ok(hipMalloc(&g_msw_hist, (size_t)W * vbf), "multishift residual window")) return 0;
            g_msw_hist_bytes = (size_t)W * vbf;
        }
        hist = (float2*)g_msw_hist;
        if (!g_msw && !hip_ok(hipMalloc(&g_msw, sizeof(MsWin)), "multishift window")) return 0;
        hipLaunchKernelGGL(k_msw_reset, dim3(1), dim3(kBlock), 0, S(), g_msw, ctl_of(kAlways));
    }
    // Reliable updates are taken at every kMsxRuEvery-th iteration (when the
    // FP32 residual has fallen by mixed_delta since the last one): the other
    // iterations run the plain CG-M finalizers and launch nothing for the
    // update.  The window continues across an update: its residuals are the
    // ones the recurrence used, replaced or not.
    constexpr int kMsxRuEvery = 8;
    auto iteration = [&](int j) -> bool {
        const bool ru_slot = j % kMsxRuEvery == 0;
        if (!launch_mdagm<float>(uf, phf, p, h, sv, m2, 2, ru_slot ? kFinMxmAlpha : kFinMsAlpha, ctl_of(kIfRunning)))
            return false;
        hipLaunchKernelGGL(k_msx_r, dim3(nbh), dim3(kBlock), 0, S(), (float2*)r, (const float2*)sv,
                           (const float2*)p, (float2*)x0f, (int)sizeh, ctl_of(kIfRunning));
        if (ru_slot) {
            // reliable update: accumulators into the FP64 solutions, r = b - M x0 in FP64
            hipLaunchKernelGGL(k_msx_flush, dim3(nb3), dim3(kBlock), 0, S(), fl, n3, ctl_of(kIfRunningRU));
            if (!launch_mdagm<double>(u, ph, x0, h64, s64, m2, 1, kFinStore, ctl_of(kIfRunningRU))) return false;
            hipLaunchKernelGGL(k_msx_resid, dim3(nbh), dim3(kBlock), 0, S(), (const double2*)db,
                               (const double2*)s64, (float2*)r, (int)sizeh, ctl_of(kIfRunningRU));
        }
        if (!launch_finalize(ru_slot ? kFinMxmLambda : kFinMsLambda, ctl_of(kIfRunning), nbh)) return false;
        if (!hist) {
            hipLaunchKernelGGL((k_ms_update<float2>), dim3(nb3), dim3(kBlock), 0, S(), (const float2*)r, (float2*)p,
                               (float2*)xsf, (float2*)pk, n3, n3, ctl_of(kIfIter, j));
            return check_launch("mixed multishift iteration");
        }
        hipLaunchKernelGGL((k_msw_p<float2>), dim3(nb3), dim3(kBlock), 0, S(), (float2*)p, (const float2*)r, n3,
                           g_msw, ctl_of(kIfIter, j), hist + (size_t)((j - 1) % W) * n3);
        if (j % W == 0) {
            // only when iteration j ran: a window cut short is flushed once, at the end
            hipLaunchKernelGGL((k_msw_flush<double2, float2>), dim3(nb3), dim3(kBlock), 0, S(), (double2*)dout,
                               (float2*)pk, (const float2*)hist, n3, order, (const MsWin*)g_msw, ctl_of(kIfIter, j));
            hipLaunchKernelGGL(k_msw_reset, dim3(1), dim3(kBlock), 0, S(), g_msw, ctl_of(kIfIter, j));
        }
        return check_launch("mixed multishift iteration");
    };
    if (!run_batches(iteration, 4, dist_ms)) return 0;
    const int cg = H().cg;
    if (cg_multishift) *cg_multishift = cg;
    if (hist)
        hipLaunchKernelGGL((k_msw_flush<double2, float2>), dim3(nb3), dim3(kBlock), 0, S(), (double2*)dout,
                           (float2*)pk, (const float2*)hist, n3, order, (const MsWin*)g_msw, ctl_of(kAlways));
    hipLaunchKernelGGL(k_msx_flush, dim3(nb3), dim3(kBlock), 0, S(), fl, n3, ctl_of(kAlways));
    if (!check_launch("flush")) return 0;
    if (verbosity_lv > 1 && !hip_ok(hipStreamSynchronize(S()), "timing")) return 0;
    const auto t_cgm = std::chrono::steady_clock::now();

    // every shift goes to the batched solver, which starts from the CG-M
    // solution with the FP64 residual b - (M + shift_k) x_k (one that is already
    // below res stops after its first inner iteration and reliable update)
    std::vector<ostp_batch_system> sys(order);
    for (int k = 0; k < order; k++) {
        ostp_batch_system& q = sys[k];
        for (int d = 0; d < 4; d++) q.nd[d] = L.nd[d];
        q.par_temp = L.par_temp;
        q.u = ip.u;
        q.u_f = ip.u_f;
        q.phases = pars->phases;
        q.phases_f = pars->phases_f;
        q.in = in;
        q.out = (vec3_soa*)((char*)out + k * vb);
        q.mass = std::sqrt(m2 + approx->RA_b[k]);
    }
    const int converged = batched_solve_internal(order, sys.data(), res, max_cg, 0.1, g_internal_ipg, cg_refine, nullptr,
                                                 true_res, half_inner != 0);
    if (verbosity_lv > 1 && hip_ok(hipStreamSynchronize(S()), "timing")) {
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_cgm).count();
        const double sec0 = std::chrono::duration<double>(t_cgm - t_start).count();
        int rmax = 0;
        for (int k = 0; cg_refine && k < order; k++) rmax = std::max(rmax, cg_refine[k]);
        printf("mixed multishift: m %g, CG-M %d iterations %.4f s, refinement of %d shifts (at most %d iterations) "
               "%.4f s\n", pars->ferm_mass, cg, sec0, order, rmax, sec);
    }
    if (verbosity_lv > 0)
        printf("Terminated mixed multishift: %d CG-M iterations, %d of %d shifts converged\n", cg, converged, order);
    return converged;
}

// Power iterations with the operator coef - Deo Doe (find_min_max.c:21-97):
// scalars on the device, iterations queued in batches like the CG.  Returns
// the last norm (NaN on failure), the iteration count in *iters.
static double power_method(const void* u, const void* ph, double coef, void* r, void* h, void* p, int* iters) {
    const long n3 = 3 * ctx().lat.sizeh;
    double sn = 0.0;
    if (!reduce_to_host(0, p, nullptr, 0, &sn, nullptr)) return NAN;
    init_state(0.0, 0.0, 1000000, 0);  // the reference has no iteration limit
    H().lambda = std::sqrt(sn);
    if (!push_state()) return NAN;
    auto iteration = [&](int) -> bool {
        hipLaunchKernelGGL(k_pm_scale, dim3(blocks_for(n3)), dim3(kBlock), 0, S(), (const double2*)p, (double2*)r, n3,
                           ctl_of(kIfRunning));
        if (!launch_mdagm<double>(u, ph, r, h, p, coef, 1, kFinStore, ctl_of(kIfRunning))) return false;  // p = coef r - Deo Doe r
        return launch_norm2<double2, kFinPower>(p, ctl_of(kIfRunning)) && check_launch("power method");
    };
    if (!run_batches(iteration)) return NAN;
    if (iters) *iters = H().cg;
    return H().lambda;
}

static bool power_buffers(const char* nm, const su3_soa* u, ferm_param* pars, vec3_soa* loc_r, vec3_soa* loc_h,
                          vec3_soa* loc_p, const void** du, const void** dph, void** dr, void** dh, void** dp) {
    const Lattice& L = ctx().lat;
    *du = resolve(u, 8 * L.su3_bytes(), nm);
    *dph = resolve(pars->phases, 8 * L.dsoa_bytes(), nm);
    *dr = resolve(loc_r, L.vec3_bytes(), nm);
    *dh = resolve(loc_h, L.vec3_bytes(), nm);
    *dp = resolve(loc_p, L.vec3_bytes(), nm);
    return *du && *dph && *dr && *dh && *dp;
}

double ostp_find_max_eigenvalue(const su3_soa* u, ferm_param* pars, vec3_soa* loc_r, vec3_soa* loc_h,
                                vec3_soa* loc_p, int* iterations) {
    const char* nm = "ker_find_max_eigenvalue_openacc";
    if (!require_init(nm)) return NAN;
    const void *du, *dph;
    void *dr, *dh, *dp;
    if (!power_buffers(nm, u, pars, loc_r, loc_h, loc_p, &du, &dph, &dr, &dh, &dp)) return NAN;
    // fermion_matrix_multiplication: m^2 in - Deo Doe in
    return power_method(du, dph, pars->ferm_mass * pars->ferm_mass, dr, dh, dp, iterations);
}

double ostp_find_min_eigenvalue(const su3_soa* u, ferm_param* pars, vec3_soa* loc_r, vec3_soa* loc_h,
                                vec3_soa* loc_p, double max, int* iterations) {
    const char* nm = "ker_find_min_eigenvalue_openacc";
    if (!require_init(nm)) return NAN;
    const void *du, *dph;
    void *dr, *dh, *dp;
    if (!power_buffers(nm, u, pars, loc_r, loc_h, loc_p, &du, &dph, &dr, &dh, &dp)) return NAN;
    // fermion_matrix_multiplication_shifted with shift = delta - m^2, delta = max - m^2
    const double m2 = pars->ferm_mass * pars->ferm_mass;
    const double delta = max - m2;
    const double norm = power_method(du, dph, m2 + (delta - m2), dr, dh, dp, iterations);
    return max - norm;
}

double ker_find_max_eigenvalue_openacc(const su3_soa* u, ferm_param* pars, vec3_soa* loc_r, vec3_soa* loc_h,
                                       vec3_soa* loc_p) {
    return ostp_find_max_eigenvalue(u, pars, loc_r, loc_h, loc_p, nullptr);
}

double ker_find_min_eigenvalue_openacc(const su3_soa* u, ferm_param* pars, vec3_soa* loc_r, vec3_soa* loc_h,
                                       vec3_soa* loc_p, double max) {
    return ostp_find_min_eigenvalue(u, pars, loc_r, loc_h, loc_p, max, nullptr);
}

// find_min_max.c:100-121: min is m^2 (the power iteration for the minimum is
// left unused there), max from the power method on loc_p1
void find_min_max_eigenvalue_soloopenacc(const su3_soa* u, ferm_param* pars, vec3_soa* loc_r, vec3_soa* loc_h,
                                         vec3_soa* loc_p1, vec3_soa* loc_p2, double* minmax) {
    (void)loc_p2;
    minmax[0] = pars->ferm_mass * pars->ferm_mass;
    minmax[1] = ker_find_max_eigenvalue_openacc(u, pars, loc_r, loc_h, loc_p1);
    if (verbosity_lv > 3)
        printf("    Computed the min and max eigs : [ min= %.18lf; max= %.18lf ]  \n", minmax[0], minmax[1]);
}

// Extended entry points returning the diagnostics the reference only prints.
int ostp_cg_solve(const su3_soa* u, ferm_param* pars, vec3_soa* solution, const vec3_soa* in, double res,
                  vec3_soa* loc_r, vec3_soa* loc_h, vec3_soa* loc_s, vec3_soa* loc_p, int max_cg, double shift,
                  int restarting_every, int* cg_return, double* true_res) {
    const char* nm = "ostp_cg_solve";
    if (!require_init(nm)) return INVERTER_FAILURE;
    const Lattice& L = ctx().lat;
    const size_t vb = L.vec3_bytes();
    CgBufs B;
    B.u = resolve(u, 8 * L.su3_bytes(), nm);
    B.ph = resolve(pars->phases, 8 * L.dsoa_bytes(), nm);
    B.b = resolve(in, vb, nm);
    B.x = resolve(solution, vb, nm);
    B.r = resolve(loc_r, vb, nm);
    B.h = resolve(loc_h, vb, nm);
    B.s = resolve(loc_s, vb, nm);
    B.p = resolve(loc_p, vb, nm);
    if (!B.u || !B.ph || !B.b || !B.x || !B.r || !B.h || !B.s || !B.p) return INVERTER_FAILURE;
    return cg_solve(B, pars->ferm_mass, shift, res, max_cg, restarting_every, cg_return, true_res);
}

int ostp_mixed_solve(inverter_package ip, ferm_param* pars, vec3_soa* solution, const vec3_soa* in, double res,
                     int max_cg, double shift, double mixed_delta, int* cg_return, int* reliable_updates,
                     double* true_res) {
    const char* nm = "ostp_mixed_solve";
    if (!require_init(nm)) return INVERTER_FAILURE;
    const Lattice& L = ctx().lat;
    const size_t vb = L.vec3_bytes(), vbf = L.vec3f_bytes();
    MxBufs B;
    B.u = resolve(ip.u, 8 * L.su3_bytes(), nm);
    B.u_f = resolve(ip.u_f, 8 * L.su3f_bytes(), nm);
    B.ph = resolve(pars->phases, 8 * L.dsoa_bytes(), nm);
    B.ph_f = resolve(pars->phases_f, 8 * L.fsoa_bytes(), nm);
    B.b = resolve(in, vb, nm);
    B.x = resolve(solution, vb, nm);
    B.r = resolve(ip.loc_r, vb, nm);
    B.h = resolve(ip.loc_h, vb, nm);
    B.s = resolve(ip.loc_s, vb, nm);
    B.rf = resolve(ip.loc_r_f, vbf, nm);
    B.hf = resolve(ip.loc_h_f, vbf, nm);
    B.sf = resolve(ip.loc_s_f, vbf, nm);
    B.pf = resolve(ip.loc_p_f, vbf, nm);
    B.outf = resolve(ip.out_f, vbf, nm);
    if (!B.u || !B.u_f || !B.ph || !B.ph_f || !B.b || !B.x || !B.r || !B.h || !B.s || !B.rf || !B.hf || !B.sf ||
        !B.pf || !B.outf)
        return INVERTER_FAILURE;
    return mixed_solve(B, pars->ferm_mass, shift, res, max_cg, mixed_delta, cg_return, reliable_updates, true_res);
}

int ostp_set_multishift_window(int w) {
    if (w < 0 || w > kMsWinMax) return fail("ostp_set_multishift_window: %d outside 0..%d", w, kMsWinMax) ? 0 : 1;
    const int prev = g_msw_forced;
    g_msw_forced = w;
    return prev;
}

int ostp_multishift_solve(const su3_soa* u, ferm_param* pars, RationalApprox* approx, vec3_soa* out,
                          const vec3_soa* in, double residuo, vec3_soa* loc_r, vec3_soa* loc_h, vec3_soa* loc_s,
                          vec3_soa* loc_p, vec3_soa* shiftferm, int max_cg, int single, int* cg_return,
                          double* shift_res) {
    const char* nm = "ostp_multishift_solve";
    if (!require_init(nm)) return INVERTER_FAILURE;
    const int order = approx->approx_order;
    if (order < 0 || order > kMaxShifts) return fail("%s: approx_order %d", nm, order) ? 0 : INVERTER_FAILURE;
    const Lattice& L = ctx().lat;
    const size_t vb = single ? L.vec3f_bytes() : L.vec3_bytes();
    const size_t nv = order > 0 ? (size_t)order : 1;
    MsBufs B;
    B.u = resolve(u, 8 * (single ? L.su3f_bytes() : L.su3_bytes()), nm);
    B.ph = resolve(single ? (const void*)pars->phases_f : (const void*)pars->phases,
                   8 * (single ? L.fsoa_bytes() : L.dsoa_bytes()), nm);
    B.b = resolve(in, vb, nm);
    B.out = resolve(out, nv * vb, nm);
    B.shiftferm = resolve(shiftferm, nv * vb, nm);
    B.r = resolve(loc_r, vb, nm);
    B.h = resolve(loc_h, vb, nm);
    B.s = resolve(loc_s, vb, nm);
    B.p = resolve(loc_p, vb, nm);
    if (!B.u || !B.ph || !B.b || !B.out || !B.shiftferm || !B.r || !B.h || !B.s || !B.p) return INVERTER_FAILURE;
    return single ? multishift_solve<float>(B, pars->ferm_mass, order, approx->RA_b, residuo, max_cg, cg_return,
                                            shift_res)
                  : multishift_solve<double>(B, pars->ferm_mass, order, approx->RA_b, residuo, max_cg, cg_return,
                                             shift_res);
}

}  // extern "C"
