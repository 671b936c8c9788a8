// Deterministic grid reductions with an on-device scalar epilogue.
//
// A reducing kernel sums its values per block with a fixed shuffle/LDS tree
// and stores the block's partial with a plain store; k_finalize, launched
// right after it on the same stream, sums the partials in block-index order
// (so the result does not depend on scheduling) and runs the solver's scalar
// update ("finalizer") on the device.  The kernel boundary orders the partial
// stores before the finalizer's loads, so no block issues a device-scope
// release or atomic (on a multi-XCD part each such release writes back the
// XCD's L2, which costs more than the reduction itself).  The solver
// recurrences never wait for a host round trip.
#pragma once

#include "ostp_runtime.h"

#include <math.h>

namespace ostp {

enum Fin : int {
    kFinStore = 0,      // st->sum = total (stand-alone reduction)
    kFinCgAlpha,        // alpha, omega = delta/alpha          (inverter_full.c:76-80)
    kFinCgLambda,       // lambda, gammag, delta, counters, exit test (inverter_full.c:86-105)
    kFinCgRestart,      // delta = |r|^2 at a (re)start          (inverter_full.c:60-66)
    kFinMsAlpha,        // CG-M alpha + shift recurrences     (inverter_multishift_full.c:116-136)
    kFinMsLambda,       // CG-M lambda, gammas, flags          (inverter_multishift_full.c:141-171)
    kFinPower,          // power method: norm = sqrt(|p|^2), stop at relative change <= 1e-5
                        //   (find_min_max.c:21-58)
    kFinMxmAlpha,       // the same two for the FP32 CG-M with FP64 reliable updates: also decide
    kFinMxmLambda,      //   whether the base residual is replaced (as kFinMxAlpha / kFinMxLambda)
    kFinMxAlpha,        // mixed: omega + reliable-update test (inverter_mixedp.c:103-114)
    kFinMxLambda,       // mixed: lambda, gammag, exit test     (inverter_mixedp.c:137-155)
    kFinStore2,         // st->sum, st->sum_im (complex dot)
    kFinFcgAlpha,       // flexible CG: alpha = <z,r> / <p,q>
    kFinFcgRes,         // flexible CG: |r|^2, counters, exit test
    kFinFcgBeta,        // flexible CG: beta = -alpha <z,q> / <z,r>_old ; <z,r> (two values)
    kFinFcgInit,        // flexible CG start: <z,r> (second value), running
    // FP16-storage inner CG with FP64 reliable updates; inner vectors carry a
    // scale so that their entries stay O(1) (rescaled at every reliable update)
    kFinHScale,         // scale = sqrt(n / |r|^2) from the FP64 residual, rho = new / old
    kFinHInit,          // delta = |r_h|^2, running
    kFinHAlpha,         // as kFinMxAlpha, rho = 1; a due reliable update stays due until done
                        // (the end of an FP16 iteration -- |r|^2, gammag, exit test -- runs
                        // inside the next D_oe kernel)
    kFinHRu,            // after a reliable update: gammag from the FP64 residual, exit test
};

// Predicates evaluated by every block at kernel entry (uniform).
enum Pred : int {
    kAlways = 0,
    kIfRunning = 1,      // status == kRunning
    kIfIter = 2,         // cg == iteration number this launch belongs to
    kIfRunningRU = 3,    // running and reliable update scheduled
    kIfRunningNoRU = 4,  // running and no reliable update scheduled
    kIfGoing = 5,        // running and not holding for a reliable update
};

struct SolverCtl {
    SolverState* st = nullptr;
    double* partials = nullptr;
    int pred = kAlways;
    int iter = 0;
};

__device__ __forceinline__ bool pred_pass(const SolverCtl& c) {
    if (c.pred == kAlways) return true;
    const SolverState* st = c.st;
    if (c.pred == kIfRunning) return st->status == kRunning;
    if (c.pred == kIfIter) return st->cg == c.iter;
    if (c.pred == kIfRunningRU) return st->status == kRunning && st->ru != 0;
    if (c.pred == kIfRunningNoRU) return st->status == kRunning && st->ru == 0;
    if (c.pred == kIfGoing) return st->status == kRunning && st->hold == 0;
    return false;
}

template <int F>
__device__ __forceinline__ void finalize(SolverState* st, double tot, double tot_im) {
    if (F == kFinStore) {
        st->sum = tot;
    } else if (F == kFinStore2) {
        st->sum = tot;
        st->sum_im = tot_im;
    } else if (F == kFinCgAlpha) {
        st->alpha = tot;
        st->omega = st->delta / tot;
    } else if (F == kFinCgRestart) {
        st->delta = tot;
        st->cg_restarted = 0;
        st->status = kRunning;
    } else if (F == kFinCgLambda) {
        const double lambda = tot;
        st->lambda = lambda;
        st->gammag = lambda / st->delta;
        st->delta = lambda;
        st->cg += 1;
        st->cg_restarted += 1;
        const bool cont = (sqrt(lambda / st->source_norm) > st->res * st->safety) &&
                          st->cg_restarted < st->restart_every;
        if (!cont) st->status = kInnerExit;
    } else if (F == kFinMsAlpha || F == kFinMxmAlpha) {
        // flags/maxiter computed at the end of the previous iteration become current
        st->flag_cur ^= 1;
        const int cur = st->flag_cur;
        st->maxiter = st->maxiter_next;
        st->alpha = tot;
        st->omega_save = st->omega;
        st->omega = -st->delta / tot;
        const double omega = st->omega, omega_save = st->omega_save, gammag = st->gammag;
        for (int k = 0; k < st->maxiter; k++) {
            if (st->flags[cur][k] == 1) {
                const double zi = st->zeta_i[k], zii = st->zeta_ii[k];
                const double ziii =
                    (zi * zii * omega_save) / (omega * gammag * (zi - zii) + zi * omega_save * (1.0 - (st->shifts[k]) * omega));
                st->zeta_iii[k] = ziii;
                st->omegas[k] = omega * ziii / zii;
            }
        }
        if (F == kFinMxmAlpha) {
            if (st->last_max_res < st->delta) st->last_max_res = st->delta;
            st->ru = (st->delta < st->mixed_delta * st->last_max_res) ? 1 : 0;
        }
    } else if (F == kFinMsLambda || F == kFinMxmLambda) {
        const double lambda = tot;
        const int cur = st->flag_cur, nxt = cur ^ 1;
        st->lambda = lambda;
        const double gammag = lambda / st->delta;
        st->gammag = gammag;
        for (int k = 0; k < st->order; k++)
            if (st->flags[cur][k] == 1)
                st->gammas[k] = gammag * st->zeta_iii[k] * st->omegas[k] / (st->zeta_ii[k] * st->omega);
        int maxiter = 0;
        const double target = st->res * st->safety;
        for (int k = 0; k < st->order; k++) {
            int f = st->flags[cur][k];
            if (f == 1) {
                const double fact = sqrt(st->delta * st->zeta_ii[k] * st->zeta_ii[k] / st->source_norm);
                if (fact < target) f = 0;
                else maxiter = k + 1;
                st->zeta_i[k] = st->zeta_ii[k];
                st->zeta_ii[k] = st->zeta_iii[k];
            }
            st->flags[nxt][k] = f;
        }
        st->maxiter_next = maxiter;
        st->delta = lambda;
        st->cg += 1;
        if (!(maxiter > 0 && st->cg < st->max_cg)) st->status = kDone;
        if (F == kFinMxmLambda && st->ru) {
            st->last_max_res = 0;
            st->ru_count += 1;
        }
    } else if (F == kFinPower) {
        // lambda holds the previous norm
        const double norm = sqrt(tot);
        double change = fabs(st->lambda - norm);
        change /= norm;
        st->lambda = norm;
        st->cg += 1;
        if (!(change > 1.0e-5) || st->cg >= st->max_cg) st->status = kDone;
    } else if (F == kFinFcgAlpha) {
        st->alpha = st->delta / tot;
        st->omega = st->alpha;  // the x/r update kernel reads omega
    } else if (F == kFinFcgRes) {
        st->lambda = tot;
        st->cg += 1;
        const bool cont = (sqrt(tot / st->source_norm) > st->res * st->safety) && st->cg < st->max_cg;
        if (!cont) st->status = kDone;
    } else if (F == kFinFcgBeta) {
        st->gammag = -st->alpha * tot / st->delta;
        st->delta = tot_im;
    } else if (F == kFinFcgInit) {
        st->delta = tot_im;
        st->status = kRunning;
    } else if (F == kFinHScale) {
        // tot: |b - M x|^2 of the FP64 residual
        st->sum = tot;
        const double n = st->nentries;
        const double sc = tot > 0.0 ? sqrt(n / tot) : 1.0;
        st->rho = st->scale > 0.0 ? sc / st->scale : 1.0;
        st->scale = sc;
    } else if (F == kFinHInit) {
        st->delta = tot;
        st->rho = 1.0;
        st->status = kRunning;
        // a start already within the tolerance (a refinement of a converged
        // solution) runs no iteration; its FP64 residual is the final one
        if (!(sqrt(st->sum / st->source_norm) > st->res * st->safety)) {
            st->status = kDone;
            st->resid_ok = 1;
        }
    } else if (F == kFinHAlpha) {
        st->alpha = tot;
        st->omega = st->delta / tot;
        st->rho = 1.0;
        if (st->last_max_res < st->delta) st->last_max_res = st->delta;
        if (st->delta < st->mixed_delta * st->last_max_res) st->ru = 1;
        // FP16 range: once |r|^2 has fallen by 1e-6 since the last update
        // (entries ~1000x below the scale they were stored at, still normal
        // FP16 numbers), the inner iterations stop until the reliable update
        // at the end of the queued block, however many iterations it holds
        if (st->delta < 1e-6 * st->last_max_res) st->hold = 1;
    } else if (F == kFinHRu) {
        // r was replaced by the (rescaled) FP64 residual and p multiplied by
        // rho: the last gammag uses the new |r|^2 over the previous |r|^2 in
        // the new scale.  If the last x/r update's |r|^2 was not consumed yet
        // (xr_fresh == 1; 2 means consumed), the previous |r|^2 is delta and
        // this ends that iteration.
        const double lambda = tot;
        st->lambda = lambda;
        double prev = st->delta_old;
        if (st->xr_fresh == 1) {
            prev = st->delta;
            st->cg += 1;
        }
        st->gammag = lambda / (prev * st->rho * st->rho);
        st->delta = lambda;
        st->ru = 0;
        st->hold = 0;
        st->xr_fresh = 0;
        st->last_max_res = 0;
        st->ru_count += 1;
        const double true2 = lambda / (st->scale * st->scale);
        if (!(sqrt(true2 / st->source_norm) > st->res * st->safety) || st->cg >= st->max_cg) {
            // x was flushed by this update and no iteration follows: the
            // FP64 residual of kFinHScale (in sum) is the final one
            st->status = kDone;
            st->resid_ok = 1;
        }
    } else if (F == kFinMxAlpha) {
        st->alpha = tot;
        st->omega = st->delta / tot;
        if (st->last_max_res < st->delta) st->last_max_res = st->delta;
        st->ru = (st->delta < st->mixed_delta * st->last_max_res) ? 1 : 0;
    } else if (F == kFinMxLambda) {
        const double lambda = tot;
        st->lambda = lambda;
        st->gammag = lambda / st->delta;
        st->delta = lambda;
        st->cg += 1;
        if (st->ru) {
            st->last_max_res = 0;
            st->ru_count += 1;
        }
        const bool cont = (sqrt(lambda / st->source_norm) > st->res * st->safety) && st->cg < st->max_cg;
        if (!cont) st->status = kDone;
    }
}

// Block partial of a grid reduction, stored at partials[2*blockIdx.x (+1)].
// All threads of the block must call this (idle threads contribute 0).
// NV = 1 or 2 values per thread.
template <int NV = 1>
__device__ __forceinline__ void block_partial(double v0, double v1, const SolverCtl& c) {
    __shared__ double lds[2][kBlock / 64];
    const double b0 = block_sum<kBlock>(v0, lds[0]);
    double b1 = 0.0;
    if (NV == 2) b1 = block_sum<kBlock>(v1, lds[1]);
    if (threadIdx.x == 0) {
        c.partials[2 * blockIdx.x] = b0;
        if (NV == 2) c.partials[2 * blockIdx.x + 1] = b1;
    }
}

// The same for a block of BLK threads with an explicit partial slot (batched
// kernels, whose hardware workgroup id is remapped to a logical tile of one
// system).
template <int BLK, int NV = 1>
__device__ __forceinline__ void block_partial_at(double v0, double v1, double* partials, int slot) {
    __shared__ double lds[2][BLK / 64];
    const double b0 = block_sum<BLK>(v0, lds[0]);
    double b1 = 0.0;
    if (NV == 2) b1 = block_sum<BLK>(v1, lds[1]);
    if (threadIdx.x == 0) {
        partials[2 * slot] = b0;
        if (NV == 2) partials[2 * slot + 1] = b1;
    }
}

// The CG-M finalizers (kFinMsAlpha / kFinMsLambda and their FP32 twins) with
// lane k of wave 0 running the recurrence of shift k: the shifts are
// independent given the base scalars, and the expressions are those of
// finalize<F>, so the results are bit-identical to its serial loops (which
// wait on device memory once per shift).  Every lane reads the old scalars
// before lane 0 stores the new ones (one wave, program order; the per-shift
// arrays and the scalars do not overlap).  maxiter (highest active shift + 1)
// is a ballot.  Needs kMaxShifts <= 64.
static_assert(kMaxShifts <= 64, "one lane per shift");
template <int F>
__device__ __forceinline__ void finalize_ms_lanes(SolverState* st, double tot, int k) {
    if (F == kFinMsAlpha || F == kFinMxmAlpha) {
        const int cur = st->flag_cur ^ 1;
        const int maxiter = st->maxiter_next;
        const double omega_save = st->omega;
        const double omega = -st->delta / tot;
        const double gammag = st->gammag;
        if (k < maxiter && st->flags[cur][k] == 1) {
            const double zi = st->zeta_i[k], zii = st->zeta_ii[k];
            const double ziii =
                (zi * zii * omega_save) / (omega * gammag * (zi - zii) + zi * omega_save * (1.0 - (st->shifts[k]) * omega));
            st->zeta_iii[k] = ziii;
            st->omegas[k] = omega * ziii / zii;
        }
        if (k == 0) {
            st->flag_cur = cur;
            st->maxiter = maxiter;
            st->alpha = tot;
            st->omega_save = omega_save;
            st->omega = omega;
            if (F == kFinMxmAlpha) {
                if (st->last_max_res < st->delta) st->last_max_res = st->delta;
                st->ru = (st->delta < st->mixed_delta * st->last_max_res) ? 1 : 0;
            }
        }
    } else {
        const double lambda = tot;
        const int cur = st->flag_cur, nxt = cur ^ 1;
        const double delta = st->delta, omega = st->omega;
        const double gammag = lambda / delta;
        const double target = st->res * st->safety;
        const double src = st->source_norm;
        const int order = st->order;
        int f = 0;
        if (k < order) {
            f = st->flags[cur][k];
            if (f == 1) {
                st->gammas[k] = gammag * st->zeta_iii[k] * st->omegas[k] / (st->zeta_ii[k] * omega);
                const double fact = sqrt(delta * st->zeta_ii[k] * st->zeta_ii[k] / src);
                if (fact < target) f = 0;
                st->zeta_i[k] = st->zeta_ii[k];
                st->zeta_ii[k] = st->zeta_iii[k];
            }
            st->flags[nxt][k] = f;
        }
        const unsigned long long act = __ballot(k < order && f == 1);
        const int maxiter = act ? 64 - __clzll(act) : 0;
        if (k == 0) {
            st->lambda = lambda;
            st->gammag = gammag;
            st->maxiter_next = maxiter;
            st->delta = lambda;
            st->cg += 1;
            if (!(maxiter > 0 && st->cg < st->max_cg)) st->status = kDone;
            if (F == kFinMxmLambda && st->ru) {
                st->last_max_res = 0;
                st->ru_count += 1;
            }
        }
    }
}

// Sum `nparts` block partials in block order and apply finalizer F.  One
// block; launched on the stream of the reducing kernel, with its SolverCtl
// (the predicate reads state that only finalizers write, so it evaluates as
// it did for the reducing kernel).
template <int F, int NV>
__global__ __launch_bounds__(kBlock) void k_finalize(SolverCtl c, int nparts) {
    if (!pred_pass(c)) return;
    __shared__ double lds[2][kBlock / 64];
    double a0 = 0.0, a1 = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kBlock) {
        a0 += c.partials[2 * i];
        if (NV == 2) a1 += c.partials[2 * i + 1];
    }
    constexpr bool kMs = F == kFinMsAlpha || F == kFinMxmAlpha || F == kFinMsLambda || F == kFinMxmLambda;
    if constexpr (kMs) {
        const double t0 = block_sum_all<kBlock>(a0, lds[0]);
        if (threadIdx.x < 64) finalize_ms_lanes<F>(c.st, t0, (int)threadIdx.x);
    } else {
        const double t0 = block_sum<kBlock>(a0, lds[0]);
        double t1 = 0.0;
        if (NV == 2) t1 = block_sum<kBlock>(a1, lds[1]);
        if (threadIdx.x == 0) finalize<F>(c.st, t0, t1);
    }
}

// Host: queue k_finalize<fin> for a reduction over `nparts` blocks.
bool launch_finalize(int fin, const SolverCtl& c, int nparts);

}  // namespace ostp
