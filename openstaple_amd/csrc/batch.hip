// Batched even-odd solves: many right-hand sides and lattices in one HIP graph.
//
// Each system is the mixed-precision CG of inverter_mixed_precision
// (OpenAcc/inverter_mixedp.c:41-181): FP32 CG on m^2 - D_eo D_oe with FP64
// reliable updates of the true residual when delta < mixed_delta * max delta
// (SAFETY 0.9, exit on sqrt(lambda/|b|^2) <= 0.9 res, final FP64 true residual).
// The reference solves one system per call; here S systems -- noise sources
// on one gauge field, parallel-tempering replicas on their own fields, and
// lattices of different extents -- advance together:
//
//  * every kernel is batched over systems (blockIdx.y = system); a system
//    whose lattice needs fewer tiles than the grid has exits those tiles;
//  * each system carries its own device-resident SolverState; kernels and
//    scalar epilogues are predicated on that state (running / reliable
//    update due), never on a host-side iteration count, so a captured
//    sequence of iterations stays valid however often it is replayed;
//  * `iters_per_graph` iterations are captured once into a HIP graph, which
//    is replayed until every system has left the running state; the host
//    reads the S states once per replay;
//  * the Dslash tile (workgroup size, t-fastest or contiguous tile order,
//    systems one after the other or interleaved so that systems on one gauge
//    field share its links through L2) is chosen per batch signature by
//    timing the candidates on the batch's own data the first time that
//    signature is seen (autotune table).
#include "ostp_dslash.h"
#include "ostp_half.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

static_assert(sizeof(ostp_batch_system) == 80 && offsetof(ostp_batch_system, mass) == 72,
              "ostp_batch_system layout");

namespace ostp {

struct BSys {
    const double2* u64[8];
    const double* ph64[8];
    const float2* u32[8];
    const float* ph32[8];
    const double2* b;
    double2* x;
    double2 *r64, *h64, *s64;
    float2 *rf, *hf, *sf, *pf, *outf;
    SiteGeom g;
    double m2, m2f;
    SolverState* st;
    double* partials;
    // FP16 inner solver: links rows 0,1 and phases (cos, sin) in half precision,
    // inner vectors in half precision (out_f stays FP32)
    const h8* ula[8];
    const h6* ulb[8];   // general phases
    const h4* ulc[8];   // phases +-1 (sign in ula)
    h6 *rh, *hh, *sh, *pph;
};

// half-sites per thread of the FP16 x/r update (kb_h_xr)
constexpr int kXrSites = 4;
// FP16 inner solver: iterations between the predicated reliable-update tails
// of a captured graph
constexpr int kRuBlock = 16;

enum BRole : int {
    kRoleF32Doe = 0,   // hf = D_oe pf                          (FP32)
    kRoleF32DeoDot,    // sf = m2f pf - D_eo hf, partial <pf,sf> (FP32)
    kRoleF64Doe,       // h64 = D_oe x                          (FP64)
    kRoleF64Deo,       // s64 = m2 x - D_eo h64                 (FP64)
    kRoleF64DeoResid,  // s64 = m2 x - D_eo h64, r64 = b - s64, partial |r64|^2 (FP64)
    kRoleH16Doe,       // hh = D_oe pph                         (FP16 storage)
    kRoleH16DeoDot,    // sh = m2f pph - D_eo hh, partial <pph,sh> (FP16 storage)
    kRoleS16Doe,       // the same two with the sign-carrying 24 B links
    kRoleS16DeoDot,
};

// p' = (half)(p gamma + r): the search-direction update of the previous
// iteration, formed where it is read (identical rounding everywhere)
__device__ __forceinline__ h6 pnext(const h6& p, const h6& r, float g) {
    h6 q;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float2 pp = h2f(p.v[c]), rr = h2f(r.v[c]);
        q.v[c] = f2h(fmaf(pp.x, g, rr.x), fmaf(pp.y, g, rr.y));
    }
    return q;
}

template <int EO, bool SGN>
__device__ __forceinline__ void dslash_site_h(const h8* const* ula, const h6* const* ulb, const h4* const* ulc,
                                              const h6* __restrict__ in, const SiteGeom& g, int idx, float2& o0,
                                              float2& o1, float2& o2) {
    const Nbrs nb = site_neighbours<EO>(g, idx);
    constexpr int F = (EO == 0) ? 0 : 1;
    constexpr int B = 1 - F;
    o0 = o1 = o2 = make_float2(0.f, 0.f);
#pragma unroll
    for (int mu = 0; mu < 4; mu++) hop_hv<true, SGN>(ula[2 * mu + B], ulb[2 * mu + B], ulc[2 * mu + B], nb.im[mu], ldg(in + nb.im[mu]), o0, o1, o2);
#pragma unroll
    for (int mu = 0; mu < 4; mu++) hop_hv<false, SGN>(ula[2 * mu + F], ulb[2 * mu + F], ulc[2 * mu + F], idx, ldg(in + nb.ip[mu]), o0, o1, o2);
    o0.x *= 0.5f; o0.y *= 0.5f; o1.x *= 0.5f; o1.y *= 0.5f; o2.x *= 0.5f; o2.y *= 0.5f;
}

__device__ __forceinline__ SolverCtl bctl(const BSys& S, int pred) {
    SolverCtl c;
    c.st = S.st;
    c.partials = S.partials;
    c.pred = pred;
    return c;
}

// logical tile of this workgroup for a system with `ntiles` tiles; -1 if idle
// Logical tile of a workgroup: XCD-contiguous ranges (each XCD walks one
// slab of every system) or, with plain, the hardware order.
__device__ __forceinline__ int tile_of(int ntiles, int plain) {
    const int blk = plain ? (int)blockIdx.x : xcd_block_id();
    return blk < ntiles ? blk : -1;
}

// (system, logical tile) of a workgroup of the (tiles, systems) grid with the
// systems fastest: the workgroups of one tile for every system follow each
// other (on one XCD, with the XCD-contiguous remap of the whole grid), so
// systems that share a gauge field read its links of that tile once from
// memory and from the XCD's L2 after that.
__device__ __forceinline__ void sys_tile_interleaved(int plain, int& sysi, int& tile) {
    const int G = (int)(gridDim.x * gridDim.y), i = (int)(blockIdx.y * gridDim.x + blockIdx.x);
    int L = i;
    if (!plain) {
        const int q = G >> 3, r = G & 7, x = i & 7, k = i >> 3;
        L = x * q + (x < r ? x : r) + k;
    }
    const int ns = (int)gridDim.y;
    sysi = L % ns;
    tile = L / ns;
}

// order: bit 0 t-fastest tiles (consecutive tiles at consecutive t), bit 1
// hardware workgroup order instead of the XCD-contiguous remap, bit 2 the
// systems interleaved (sys_tile_interleaved) instead of one after the other
template <int ROLE, int BLK>
__global__ __launch_bounds__(BLK) void kb_dslash(const BSys* __restrict__ sys, int pred, int order) {
    const int tmajor = order & 1;
    using R = typename std::conditional<(ROLE <= kRoleF32DeoDot || ROLE >= kRoleH16Doe), float, double>::type;
    using C = typename Cplx<R>::T;
    int sysi = (int)blockIdx.y, tile = -1;
    if (order & 4) sys_tile_interleaved(order & 2, sysi, tile);
    const BSys& S = sys[sysi];
    if (!pred_pass(bctl(S, pred))) return;
    const SiteGeom& g = S.g;
    const int ntiles = (g.sizeh + BLK - 1) / BLK;
    if (order & 4) {
        if (tile >= ntiles) tile = -1;
    } else {
        tile = tile_of(ntiles, order & 2);
    }
    if (tile < 0) return;
    const int s3 = g.sizeh / g.nd3;
    int idx;
    if (tmajor && s3 % BLK == 0) {
        const int q = tile / g.nd3, tt = tile - q * g.nd3;
        idx = tt * s3 + q * BLK + (int)threadIdx.x;
    } else {
        idx = tile * BLK + (int)threadIdx.x;
    }
    if constexpr (ROLE >= kRoleH16Doe) {
        constexpr int EOH = (ROLE == kRoleH16Doe || ROLE == kRoleS16Doe) ? 1 : 0;
        constexpr bool SGN = ROLE >= kRoleS16Doe;
        double dot = 0.0;
        float gm = 0.f;
        if constexpr (EOH == 1) {
            // End of the previous iteration, done here instead of in a
            // one-block kernel: every block sums the x/r update's |r|^2
            // partials (odd slots, kBlock tiles) in the same order, so every
            // block takes the same gammag and the same exit decision.  Block 0
            // publishes gammag/lambda; the D_eo kernel commits delta and the
            // counter (nothing there reads them); when the iteration stops
            // here, block 0 commits everything (the other blocks only return).
            SolverState* st = S.st;
            if (st->xr_fresh) {
                __shared__ double lds[BLK / 64];
                const int np = (g.sizeh + kBlock * kXrSites - 1) / (kBlock * kXrSites);
                double a = 0.0;
                for (int j = threadIdx.x; j < np; j += BLK) a += S.partials[2 * j + 1];
                const double lambda = block_sum_all<BLK>(a, lds);
                const double delta = st->delta;
                const int cg1 = st->cg + 1;
                const double true2 = lambda / (st->scale * st->scale);
                const bool at_max = cg1 >= st->max_cg;
                const bool conv = !(sqrt(true2 / st->source_norm) > st->res * st->safety);
                gm = (float)(lambda / delta);
                if (tile == 0 && threadIdx.x == 0) {
                    st->gammag = lambda / delta;
                    st->lambda = lambda;
                    if (at_max || conv) {
                        st->delta_old = delta;
                        st->delta = lambda;
                        st->cg = cg1;
                        st->xr_fresh = 2;  // consumed (still nonzero: the other blocks take this branch)
                        if (at_max) {
                            st->status = kDone;
                        } else {
                            st->ru = 1;  // inner residual converged: confirm with the FP64 residual
                            st->hold = 1;
                        }
                    }
                }
                if (at_max || conv) return;
            } else {
                gm = (float)st->gammag;  // first iteration after the start or a reliable update
            }
        } else {
            gm = (float)S.st->gammag;
            if (tile == 0 && threadIdx.x == 0 && S.st->xr_fresh) {
                SolverState* st = S.st;
                st->delta_old = st->delta;
                st->delta = st->lambda;
                st->cg += 1;
                st->xr_fresh = 0;
            }
        }
        if (idx < g.sizeh) {
            float2 o0, o1, o2;
            if constexpr (EOH == 1) {
                // D_oe of p' = p gammag + r (the previous iteration's direction update)
                const Nbrs nb = site_neighbours<1>(g, idx);
                o0 = o1 = o2 = make_float2(0.f, 0.f);
#pragma unroll
                for (int mu = 0; mu < 4; mu++) {
                    const int j = nb.im[mu];
                    hop_hv<true, SGN>(S.ula[2 * mu], S.ulb[2 * mu], S.ulc[2 * mu], j,
                                      pnext(ldg(S.pph + j), ldg(S.rh + j), gm), o0, o1, o2);
                }
#pragma unroll
                for (int mu = 0; mu < 4; mu++) {
                    const int j = nb.ip[mu];
                    hop_hv<false, SGN>(S.ula[2 * mu + 1], S.ulb[2 * mu + 1], S.ulc[2 * mu + 1], idx,
                                       pnext(ldg(S.pph + j), ldg(S.rh + j), gm), o0, o1, o2);
                }
                o0.x *= 0.5f; o0.y *= 0.5f; o1.x *= 0.5f; o1.y *= 0.5f; o2.x *= 0.5f; o2.y *= 0.5f;
                h6 hv;
                hv.v[0] = f2h(o0.x, o0.y);
                hv.v[1] = f2h(o1.x, o1.y);
                hv.v[2] = f2h(o2.x, o2.y);
                stg(S.hh + idx, hv);
            } else {
                dslash_site_h<0, SGN>(S.ula, S.ulb, S.ulc, S.hh, g, idx, o0, o1, o2);
                const float m2 = (float)S.m2f;
                // p' at this site, stored: the direction of this iteration
                const h6 q = pnext(ldg(S.pph + idx), ldg(S.rh + idx), gm);
                stg(S.pph + idx, q);
                const float2 x0 = h2f(q.v[0]), x1 = h2f(q.v[1]), x2 = h2f(q.v[2]);
                h6 sv;
                sv.v[0] = f2h(x0.x * m2 - o0.x, x0.y * m2 - o0.y);
                sv.v[1] = f2h(x1.x * m2 - o1.x, x1.y * m2 - o1.y);
                sv.v[2] = f2h(x2.x * m2 - o2.x, x2.y * m2 - o2.y);
                stg(S.sh + idx, sv);
                dot = re_dot(x0, h2f(sv.v[0])) + re_dot(x1, h2f(sv.v[1])) + re_dot(x2, h2f(sv.v[2]));
            }
        }
        if constexpr (EOH == 0) block_partial_at<BLK, 1>(dot, 0.0, S.partials, tile);
        return;
    } else {
    constexpr int EO = (ROLE == kRoleF32Doe || ROLE == kRoleF64Doe) ? 1 : 0;
    const C* const* u;
    const R* const* ph;
    const C* in;
    C* out;
    const C* x = nullptr;
    double m2 = 0.0;
    if constexpr (ROLE == kRoleF32Doe) {
        u = S.u32; ph = S.ph32; in = S.pf; out = S.hf;
    } else if constexpr (ROLE == kRoleF32DeoDot) {
        u = S.u32; ph = S.ph32; in = S.hf; out = S.sf; x = S.pf; m2 = S.m2f;
    } else if constexpr (ROLE == kRoleF64Doe) {
        u = S.u64; ph = S.ph64; in = S.x; out = S.h64;
    } else {
        u = S.u64; ph = S.ph64; in = S.h64; out = S.s64; x = S.x; m2 = S.m2;  // F64Deo, F64DeoResid
    }
    double dot = 0.0;
    if (idx < g.sizeh) {
        C o0, o1, o2;
        dslash_site<R, EO>(u, ph, in, g, idx, o0, o1, o2);
        if constexpr (EO == 1) {
            out[idx] = o0;
            out[g.sizeh + idx] = o1;
            out[2 * g.sizeh + idx] = o2;
        } else {
            const C x0 = x[idx], x1 = x[g.sizeh + idx], x2 = x[2 * g.sizeh + idx];
            const C s0 = mk<C>((double)x0.x * m2 - (double)o0.x, (double)x0.y * m2 - (double)o0.y);
            const C s1 = mk<C>((double)x1.x * m2 - (double)o1.x, (double)x1.y * m2 - (double)o1.y);
            const C s2 = mk<C>((double)x2.x * m2 - (double)o2.x, (double)x2.y * m2 - (double)o2.y);
            out[idx] = s0;
            out[g.sizeh + idx] = s1;
            out[2 * g.sizeh + idx] = s2;
            if constexpr (ROLE == kRoleF32DeoDot) dot = re_dot(x0, s0) + re_dot(x1, s1) + re_dot(x2, s2);
            if constexpr (ROLE == kRoleF64DeoResid) {
                const C sv[3] = {s0, s1, s2};
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const long k = (long)c * g.sizeh + idx;
                    const double2 bb = S.b[k];
                    const double2 rr = make_double2(bb.x - (double)sv[c].x, bb.y - (double)sv[c].y);
                    S.r64[k] = rr;
                    dot += rr.x * rr.x + rr.y * rr.y;
                }
            }
        }
    }
    if constexpr (ROLE == kRoleF32DeoDot || ROLE == kRoleF64DeoResid) block_partial_at<BLK, 1>(dot, 0.0, S.partials, tile);
    }
}

// Block sums of one value per system of a chunk into each active system's
// partial slot: every wave reduces all NR values, one barrier, thread 0
// adds the wave sums in wave order (the order of block_partial_at, so a
// system's partial does not depend on its chunk).  One LDS row per system:
// calling block_partial_at once per system would let a wave overwrite a
// wave sum of the previous system before thread 0 has read it.
template <int BLK, int NR>
__device__ __forceinline__ void chunk_partials(const double (&v)[NR], const bool (&act)[NR], const int (&sidx)[NR],
                                               const BSys* __restrict__ sys, int slot) {
    __shared__ double lds[NR][BLK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        double x = v[r];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) lds[r][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (!act[r]) continue;
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < BLK / 64; w++) t += lds[r][w];
            sys[sidx[r]].partials[2 * slot] = t;
        }
    }
}

// --------------------------------------- FP32/FP64 hops, several systems
// Systems of a batch that share a gauge field (noise sources on one
// configuration, the shifts of a rational solve) form chunks of up to NR; a
// chunk's workgroup computes the hops of all of them (the FP32 inner hops or
// the FP64 ones): each link (two rows, the rebuilt third, its phase) is
// loaded and prepared once per half-site and applied to the NR systems'
// spinors, with the single-system hop's arithmetic for each.  Links are 768
// of the 928 B a half-site's FP64 hop moves (416 of 512 in FP32), so the
// link traffic of the chunk's systems drops NR-fold.
template <int ROLE, int BLK, int NR>
__global__ __launch_bounds__(BLK) void kb_dslash_mr(const BSys* __restrict__ sys, const int* __restrict__ chunks,
                                                    int pred, int order) {
    static_assert(ROLE <= kRoleF64DeoResid, "FP32 and FP64 roles");
    using R = typename std::conditional<(ROLE <= kRoleF32DeoDot), float, double>::type;
    using C = typename Cplx<R>::T;
    constexpr int EO = (ROLE == kRoleF32Doe || ROLE == kRoleF64Doe) ? 1 : 0;
    int ci = (int)blockIdx.y, tile = -1;
    if (order & 4) sys_tile_interleaved(order & 2, ci, tile);
    int sidx[NR];
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        sidx[r] = chunks[NR * ci + r];
        act[r] = sidx[r] >= 0 && pred_pass(bctl(sys[sidx[r]], pred));
        any = any || act[r];
    }
    if (!any) return;
    const BSys& S0 = sys[sidx[0]];
    const SiteGeom& g = S0.g;
    const int ntiles = (g.sizeh + BLK - 1) / BLK;
    if (order & 4) {
        if (tile >= ntiles) tile = -1;
    } else {
        tile = tile_of(ntiles, order & 2);
    }
    if (tile < 0) return;
    const int s3 = g.sizeh / g.nd3;
    int idx;
    if ((order & 1) && s3 % BLK == 0) {
        const int q = tile / g.nd3, tt = tile - q * g.nd3;
        idx = tt * s3 + q * BLK + (int)threadIdx.x;
    } else {
        idx = tile * BLK + (int)threadIdx.x;
    }
    double dot[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) dot[r] = 0.0;
    if (idx < g.sizeh) {
        const C* in[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const BSys& S = sys[act[r] ? sidx[r] : sidx[0]];
            if constexpr (ROLE == kRoleF32Doe) in[r] = S.pf;
            else if constexpr (ROLE == kRoleF32DeoDot) in[r] = S.hf;
            else if constexpr (ROLE == kRoleF64Doe) in[r] = S.x;
            else in[r] = S.h64;
        }
        const C* const* u;
        const R* const* ph;
        if constexpr (sizeof(R) == 4) {
            u = S0.u32;
            ph = S0.ph32;
        } else {
            u = S0.u64;
            ph = S0.ph64;
        }
        C o[NR][3];
#pragma unroll
        for (int r = 0; r < NR; r++) o[r][0] = o[r][1] = o[r][2] = mk<C>(0.0, 0.0);
        const Nbrs nb = site_neighbours<EO>(g, idx);
        constexpr int F = EO == 0 ? 0 : 1, B = 1 - F;
#pragma unroll
        for (int mu = 0; mu < 4; mu++)
            hop_mr<R, true, NR>(u[2 * mu + B], ph[2 * mu + B], in, act, g.sizeh, nb.im[mu], nb.im[mu], o);
#pragma unroll
        for (int mu = 0; mu < 4; mu++)
            hop_mr<R, false, NR>(u[2 * mu + F], ph[2 * mu + F], in, act, g.sizeh, idx, nb.ip[mu], o);
        const R half = (R)0.5;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (!act[r]) continue;
            const BSys& S = sys[sidx[r]];
            C o0 = o[r][0], o1 = o[r][1], o2 = o[r][2];
            o0.x *= half; o0.y *= half; o1.x *= half; o1.y *= half; o2.x *= half; o2.y *= half;
            if constexpr (EO == 1) {
                C* out = ROLE == kRoleF32Doe ? (C*)S.hf : (C*)S.h64;
                out[idx] = o0;
                out[g.sizeh + idx] = o1;
                out[2 * g.sizeh + idx] = o2;
            } else {
                const C* x = ROLE == kRoleF32DeoDot ? (const C*)S.pf : (const C*)S.x;
                C* out = ROLE == kRoleF32DeoDot ? (C*)S.sf : (C*)S.s64;
                const double m2 = ROLE == kRoleF32DeoDot ? S.m2f : S.m2;
                const C x0 = x[idx], x1 = x[g.sizeh + idx], x2 = x[2 * g.sizeh + idx];
                const C sv[3] = {mk<C>((double)x0.x * m2 - (double)o0.x, (double)x0.y * m2 - (double)o0.y),
                                 mk<C>((double)x1.x * m2 - (double)o1.x, (double)x1.y * m2 - (double)o1.y),
                                 mk<C>((double)x2.x * m2 - (double)o2.x, (double)x2.y * m2 - (double)o2.y)};
#pragma unroll
                for (int c = 0; c < 3; c++) out[(long)c * g.sizeh + idx] = sv[c];
                if constexpr (ROLE == kRoleF32DeoDot) dot[r] = re_dot(x0, sv[0]) + re_dot(x1, sv[1]) + re_dot(x2, sv[2]);
                if constexpr (ROLE == kRoleF64DeoResid) {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        const long k = (long)c * g.sizeh + idx;
                        const double2 bb = S.b[k];
                        const double2 rr = make_double2(bb.x - (double)sv[c].x, bb.y - (double)sv[c].y);
                        S.r64[k] = rr;
                        dot[r] += rr.x * rr.x + rr.y * rr.y;
                    }
                }
            }
        }
    }
    if constexpr (ROLE == kRoleF32DeoDot || ROLE == kRoleF64DeoResid) chunk_partials<BLK, NR>(dot, act, sidx, sys, tile);
}

// ------------------------------------------- FP16 hops, several systems
// The inner FP16 hops (kRoleH16*/kRoleS16*) by chunks of systems on one gauge
// field, as kb_dslash_mr does for the FP64 hops: one thread decodes each FP16
// link record once per half-site and applies it to the chunk's systems, each
// with its own direction update p' = p gammag + r on the fly, its own scalar
// epilogue and its own partials, exactly as kb_dslash does for one system.
template <bool DAG, bool SGN, int NR>
__device__ __forceinline__ void hop_hv_mr(const h8* __restrict__ LA, const h6* __restrict__ LB,
                                          const h4* __restrict__ LC, int iu, const h6 (&vin)[NR],
                                          const bool (&act)[NR], float2 (&o)[NR][3]) {
    const h8 la = ldg(LA + iu);
    float2 a0 = h2f(la.v[0]);
    const float2 a1 = h2f(la.v[1]), a2 = h2f(la.v[2]), b0 = h2f(la.v[3]);
    float2 b1, b2, ph = make_float2(1.f, 0.f);
    bool neg = false;
    if constexpr (SGN) {
        const h4 lc = ldg(LC + iu);
        b1 = h2f(lc.v[0]);
        b2 = h2f(lc.v[1]);
        const unsigned short bits = __builtin_bit_cast(unsigned short, la.v[0].x);
        neg = (bits & 0x4000u) != 0;
        a0.x = (float)__builtin_bit_cast(_Float16, (unsigned short)(bits & 0xbfffu));
    } else {
        const h6 lb = ldg(LB + iu);
        b1 = h2f(lb.v[0]);
        b2 = h2f(lb.v[1]);
        const float2 cs = h2f(lb.v[2]);
        ph = DAG ? make_float2(cs.x, -cs.y) : cs;
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        if (!act[r]) continue;
        float2 v0, v1, v2;
        if constexpr (SGN) {
            const float2 i0 = h2f(vin[r].v[0]), i1 = h2f(vin[r].v[1]), i2 = h2f(vin[r].v[2]);
            v0 = neg ? make_float2(-i0.x, -i0.y) : i0;
            v1 = neg ? make_float2(-i1.x, -i1.y) : i1;
            v2 = neg ? make_float2(-i2.x, -i2.y) : i2;
        } else {
            v0 = cmul(h2f(vin[r].v[0]), ph);
            v1 = cmul(h2f(vin[r].v[1]), ph);
            v2 = cmul(h2f(vin[r].v[2]), ph);
        }
        hop_rows<DAG>(a0, a1, a2, b0, b1, b2, v0, v1, v2, o[r][0], o[r][1], o[r][2]);
    }
}

template <int ROLE, int BLK, int NR>
__global__ __launch_bounds__(BLK) void kb_dslash_h_mr(const BSys* __restrict__ sys, const int* __restrict__ chunks,
                                                      int pred, int order) {
    static_assert(ROLE >= kRoleH16Doe, "FP16 roles");
    constexpr int EOH = (ROLE == kRoleH16Doe || ROLE == kRoleS16Doe) ? 1 : 0;
    constexpr bool SGN = ROLE >= kRoleS16Doe;
    int ci = (int)blockIdx.y, tile = -1;
    if (order & 4) sys_tile_interleaved(order & 2, ci, tile);
    int sidx[NR];
    bool act[NR];
    bool any = false;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        sidx[r] = chunks[NR * ci + r];
        act[r] = sidx[r] >= 0 && pred_pass(bctl(sys[sidx[r]], pred));
        any = any || act[r];
    }
    if (!any) return;
    const BSys& S0 = sys[sidx[0]];
    const SiteGeom& g = S0.g;
    const int ntiles = (g.sizeh + BLK - 1) / BLK;
    if (order & 4) {
        if (tile >= ntiles) tile = -1;
    } else {
        tile = tile_of(ntiles, order & 2);
    }
    if (tile < 0) return;
    const int s3 = g.sizeh / g.nd3;
    int idx;
    if ((order & 1) && s3 % BLK == 0) {
        const int q = tile / g.nd3, tt = tile - q * g.nd3;
        idx = tt * s3 + q * BLK + (int)threadIdx.x;
    } else {
        idx = tile * BLK + (int)threadIdx.x;
    }
    // each system's end of the previous iteration (D_oe) or scalar commit
    // (D_eo), the single-system kernel's code with "stop" clearing act[r]
    __shared__ double lds[NR][BLK / 64];
    float gm[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        gm[r] = 0.f;
        if (!act[r]) continue;
        const BSys& S = sys[sidx[r]];
        SolverState* st = S.st;
        if constexpr (EOH == 1) {
            if (st->xr_fresh) {
                const int np = (g.sizeh + kBlock * kXrSites - 1) / (kBlock * kXrSites);
                double a = 0.0;
                for (int j = threadIdx.x; j < np; j += BLK) a += S.partials[2 * j + 1];
                const double lambda = block_sum_all<BLK>(a, lds[r]);
                const double delta = st->delta;
                const int cg1 = st->cg + 1;
                const double true2 = lambda / (st->scale * st->scale);
                const bool at_max = cg1 >= st->max_cg;
                const bool conv = !(sqrt(true2 / st->source_norm) > st->res * st->safety);
                gm[r] = (float)(lambda / delta);
                if (tile == 0 && threadIdx.x == 0) {
                    st->gammag = lambda / delta;
                    st->lambda = lambda;
                    if (at_max || conv) {
                        st->delta_old = delta;
                        st->delta = lambda;
                        st->cg = cg1;
                        st->xr_fresh = 2;
                        if (at_max) {
                            st->status = kDone;
                        } else {
                            st->ru = 1;
                            st->hold = 1;
                        }
                    }
                }
                if (at_max || conv) act[r] = false;
            } else {
                gm[r] = (float)st->gammag;
            }
        } else {
            gm[r] = (float)st->gammag;
            if (tile == 0 && threadIdx.x == 0 && st->xr_fresh) {
                st->delta_old = st->delta;
                st->delta = st->lambda;
                st->cg += 1;
                st->xr_fresh = 0;
            }
        }
    }
    any = false;
#pragma unroll
    for (int r = 0; r < NR; r++) any = any || act[r];
    if (!any) return;
    double dot[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) dot[r] = 0.0;
    if (idx < g.sizeh) {
        const BSys* Sr[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) Sr[r] = &sys[act[r] ? sidx[r] : sidx[0]];
        float2 o[NR][3];
#pragma unroll
        for (int r = 0; r < NR; r++) o[r][0] = o[r][1] = o[r][2] = make_float2(0.f, 0.f);
        const Nbrs nb = site_neighbours<EOH>(g, idx);
        constexpr int F = EOH == 0 ? 0 : 1, B = 1 - F;
        h6 vin[NR];
#pragma unroll
        for (int mu = 0; mu < 4; mu++) {
            const int j = nb.im[mu];
#pragma unroll
            for (int r = 0; r < NR; r++)
                if (act[r])
                    vin[r] = EOH == 1 ? pnext(ldg(Sr[r]->pph + j), ldg(Sr[r]->rh + j), gm[r]) : ldg(Sr[r]->hh + j);
            hop_hv_mr<true, SGN, NR>(S0.ula[2 * mu + B], S0.ulb[2 * mu + B], S0.ulc[2 * mu + B], j, vin, act, o);
        }
#pragma unroll
        for (int mu = 0; mu < 4; mu++) {
            const int j = nb.ip[mu];
#pragma unroll
            for (int r = 0; r < NR; r++)
                if (act[r])
                    vin[r] = EOH == 1 ? pnext(ldg(Sr[r]->pph + j), ldg(Sr[r]->rh + j), gm[r]) : ldg(Sr[r]->hh + j);
            hop_hv_mr<false, SGN, NR>(S0.ula[2 * mu + F], S0.ulb[2 * mu + F], S0.ulc[2 * mu + F], idx, vin, act, o);
        }
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (!act[r]) continue;
            const BSys& S = *Sr[r];
            float2 o0 = o[r][0], o1 = o[r][1], o2 = o[r][2];
            o0.x *= 0.5f; o0.y *= 0.5f; o1.x *= 0.5f; o1.y *= 0.5f; o2.x *= 0.5f; o2.y *= 0.5f;
            if constexpr (EOH == 1) {
                h6 hv;
                hv.v[0] = f2h(o0.x, o0.y);
                hv.v[1] = f2h(o1.x, o1.y);
                hv.v[2] = f2h(o2.x, o2.y);
                stg(S.hh + idx, hv);
            } else {
                const float m2 = (float)S.m2f;
                const h6 q = pnext(ldg(S.pph + idx), ldg(S.rh + idx), gm[r]);
                stg(S.pph + idx, q);
                const float2 x0 = h2f(q.v[0]), x1 = h2f(q.v[1]), x2 = h2f(q.v[2]);
                h6 sv;
                sv.v[0] = f2h(x0.x * m2 - o0.x, x0.y * m2 - o0.y);
                sv.v[1] = f2h(x1.x * m2 - o1.x, x1.y * m2 - o1.y);
                sv.v[2] = f2h(x2.x * m2 - o2.x, x2.y * m2 - o2.y);
                stg(S.sh + idx, sv);
                dot[r] = re_dot(x0, h2f(sv.v[0])) + re_dot(x1, h2f(sv.v[1])) + re_dot(x2, h2f(sv.v[2]));
            }
        }
    }
    if constexpr (EOH == 0) chunk_partials<BLK, NR>(dot, act, sidx, sys, tile);
}

// Elementwise kernels: one thread per half-site, 3 colours each, kBlock threads.
#define BATCH_SITE_PROLOGUE(pred)                                     \
    const BSys& S = sys[blockIdx.y];                                  \
    if (!pred_pass(bctl(S, pred))) return;                            \
    const int ntiles = (S.g.sizeh + kBlock - 1) / kBlock;             \
    if ((int)blockIdx.x >= ntiles) return;                            \
    const int i = blockIdx.x * kBlock + threadIdx.x;                  \
    const long sh = S.g.sizeh;

// r = b - s (FP64); r_f = (float) r; p_f = r_f; out_f = 0; partial |r_f|^2  (inverter_mixedp.c:74-88)
__global__ __launch_bounds__(kBlock) void kb_mx_init(const BSys* __restrict__ sys) {
    BATCH_SITE_PROLOGUE(kAlways)
    double v = 0.0;
    if (i < sh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = c * sh + i;
            const double2 bb = S.b[k], ss = S.s64[k];
            const double2 rr = make_double2(bb.x - ss.x, bb.y - ss.y);
            S.r64[k] = rr;
            const float2 f = make_float2((float)rr.x, (float)rr.y);
            S.rf[k] = f;
            S.pf[k] = f;
            S.outf[k] = make_float2(0.f, 0.f);
            v += (double)f.x * f.x + (double)f.y * f.y;
        }
    }
    block_partial_at<kBlock, 1>(v, 0.0, S.partials, blockIdx.x);
}

// out_f += omega p_f; then r_f -= omega s_f with partial |r_f|^2, or (reliable
// update due) x += out_f, out_f = 0   (inverter_mixedp.c:107-134)
__global__ __launch_bounds__(kBlock) void kb_mx_xr(const BSys* __restrict__ sys) {
    BATCH_SITE_PROLOGUE(kIfRunning)
    const double om = S.st->omega;
    const int ru = S.st->ru;
    double v = 0.0;
    if (i < sh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = c * sh + i;
            const float2 pp = S.pf[k], oo = S.outf[k];
            const float2 on =
                make_float2((float)((double)pp.x * om + (double)oo.x), (float)((double)pp.y * om + (double)oo.y));
            if (!ru) {
                S.outf[k] = on;
                const float2 ss = S.sf[k], rr = S.rf[k];
                const float2 rn = make_float2((float)((double)ss.x * (-om) + (double)rr.x),
                                              (float)((double)ss.y * (-om) + (double)rr.y));
                S.rf[k] = rn;
                v += (double)rn.x * rn.x + (double)rn.y * rn.y;
            } else {
                const double2 xx = S.x[k];
                S.x[k] = make_double2(xx.x + (double)on.x, xx.y + (double)on.y);
                S.outf[k] = make_float2(0.f, 0.f);
            }
        }
    }
    if (!ru) block_partial_at<kBlock, 1>(v, 0.0, S.partials, blockIdx.x);
}

// reliable update tail: r = b - s (FP64), r_f = (float) r, partial |r_f|^2  (inverter_mixedp.c:121-124,137)
__global__ __launch_bounds__(kBlock) void kb_mx_resid(const BSys* __restrict__ sys) {
    BATCH_SITE_PROLOGUE(kIfRunningRU)
    double v = 0.0;
    if (i < sh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = c * sh + i;
            const double2 bb = S.b[k], ss = S.s64[k];
            const double2 rr = make_double2(bb.x - ss.x, bb.y - ss.y);
            S.r64[k] = rr;
            const float2 f = make_float2((float)rr.x, (float)rr.y);
            S.rf[k] = f;
            v += (double)f.x * f.x + (double)f.y * f.y;
        }
    }
    block_partial_at<kBlock, 1>(v, 0.0, S.partials, blockIdx.x);
}

// p_f = p_f * gammag + r_f   (inverter_mixedp.c:150)
__global__ __launch_bounds__(kBlock) void kb_cg_p(const BSys* __restrict__ sys) {
    BATCH_SITE_PROLOGUE(kIfRunning)
    if (i >= sh) return;
    const double g = S.st->gammag;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const long k = c * sh + i;
        const float2 pp = S.pf[k], rr = S.rf[k];
        S.pf[k] = make_float2((float)((double)pp.x * g + (double)rr.x), (float)((double)pp.y * g + (double)rr.y));
    }
}

// x += (double) out_f   (inverter_mixedp.c:24-34,165)
__global__ __launch_bounds__(kBlock) void kb_add_f_to_d(const BSys* __restrict__ sys) {
    BATCH_SITE_PROLOGUE(kAlways)
    if (i >= sh) return;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const long k = c * sh + i;
        const double2 xx = S.x[k];
        const float2 ff = S.outf[k];
        S.x[k] = make_double2(xx.x + (double)ff.x, xx.y + (double)ff.y);
    }
}

// partial of |b - s|^2 (true residual, s = M x) or of |b|^2 (with_s = 0)
__global__ __launch_bounds__(kBlock) void kb_resid_norm(const BSys* __restrict__ sys, int with_s) {
    BATCH_SITE_PROLOGUE(kAlways)
    double v = 0.0;
    if (i < sh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = c * sh + i;
            double2 rr = S.b[k];
            if (with_s) {
                const double2 ss = S.s64[k];
                rr = make_double2(rr.x - ss.x, rr.y - ss.y);
            }
            v += rr.x * rr.x + rr.y * rr.y;
        }
    }
    block_partial_at<kBlock, 1>(v, 0.0, S.partials, blockIdx.x);
}

// ---- FP16-storage inner solver (FP32 arithmetic, FP64 reliable updates).
// Inner vectors hold scale x (true vector); the scale makes the residual's
// entries O(1) after every reliable update (see kFinHScale).

// r_h = (half) scale r64; start: p_h = r_h, out_f = 0; otherwise p_h *= rho.
// Partial |r_h|^2 of the stored (rounded) values.
__global__ __launch_bounds__(kBlock) void kb_h_rescale(const BSys* __restrict__ sys, int pred, int start) {
    BATCH_SITE_PROLOGUE(pred)
    const double sc = S.st->scale;
    const float rho = (float)S.st->rho;
    double v = 0.0;
    if (i < sh) {
        h6 rv, pv = start ? h6{} : ldg(S.pph + i);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = c * sh + i;
            const double2 rr = S.r64[k];
            rv.v[c] = f2h((float)(sc * rr.x), (float)(sc * rr.y));
            const float2 rf = h2f(rv.v[c]);
            v += (double)rf.x * rf.x + (double)rf.y * rf.y;
            if (start) {
                pv.v[c] = rv.v[c];
                S.outf[k] = make_float2(0.f, 0.f);
            } else {
                const float2 pp = h2f(pv.v[c]);
                pv.v[c] = f2h(pp.x * rho, pp.y * rho);
            }
        }
        stg(S.rh + i, rv);
        stg(S.pph + i, pv);
    }
    block_partial_at<kBlock, 1>(v, 0.0, S.partials, blockIdx.x);
}

// alpha from the D_eo partials <p', s> (even slots, summed by every block in
// the same order, so every block computes the same omega; block 0 commits the
// scalars), then out_f += omega p_h, r_h -= omega s_h with partial |r_h|^2 in
// the odd slots (the even ones are still being read by other blocks).  Each
// block covers kXrSites consecutive tiles of kBlock half-sites, which keeps
// the number of |r|^2 partials (read by every block of the next D_oe) small.
__global__ __launch_bounds__(kBlock) void kb_h_xr(const BSys* __restrict__ sys, int dot_tile) {
    const BSys& S = sys[blockIdx.y];
    if (!pred_pass(bctl(S, kIfGoing))) return;
    const long sh = S.g.sizeh;
    const int nxr = (int)((sh + kBlock * kXrSites - 1) / (kBlock * kXrSites));
    if ((int)blockIdx.x >= nxr) return;
    // the site data does not depend on alpha: load it before the sum so the
    // memory latency overlaps the partial reads
    h6 pv[kXrSites], sv[kXrSites], rv[kXrSites];
    float2 oo[kXrSites][3];
#pragma unroll
    for (int u = 0; u < kXrSites; u++) {
        const long i = ((long)blockIdx.x * kXrSites + u) * kBlock + threadIdx.x;
        if (i < sh) {
            pv[u] = ldg(S.pph + i);
            sv[u] = ldg(S.sh + i);
            rv[u] = ldg(S.rh + i);
#pragma unroll
            for (int c = 0; c < 3; c++) oo[u][c] = S.outf[c * sh + i];
        }
    }
    __shared__ double lds[kBlock / 64];
    const int nparts = (int)((sh + dot_tile - 1) / dot_tile);
    double a = 0.0;
    for (int j = threadIdx.x; j < nparts; j += kBlock) a += S.partials[2 * j];
    const double tot = block_sum_all<kBlock>(a, lds);
    const float om = (float)(S.st->delta / tot);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        finalize<kFinHAlpha>(S.st, tot, 0.0);
        S.st->xr_fresh = 1;  // the next D_oe kernel ends this iteration
    }
    double v = 0.0;
#pragma unroll
    for (int u = 0; u < kXrSites; u++) {
        const long i = ((long)blockIdx.x * kXrSites + u) * kBlock + threadIdx.x;
        if (i < sh) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float2 pp = h2f(pv[u].v[c]);
                S.outf[c * sh + i] = make_float2(fmaf(pp.x, om, oo[u][c].x), fmaf(pp.y, om, oo[u][c].y));
                const float2 ss = h2f(sv[u].v[c]), rr = h2f(rv[u].v[c]);
                rv[u].v[c] = f2h(fmaf(ss.x, -om, rr.x), fmaf(ss.y, -om, rr.y));
                const float2 rf = h2f(rv[u].v[c]);
                v += (double)rf.x * rf.x + (double)rf.y * rf.y;
            }
            stg(S.rh + i, rv[u]);
        }
    }
    block_partial_at<kBlock, 1>(v, 0.0, S.partials + 1, blockIdx.x);
}

// x += out_f / scale, out_f = 0
__global__ __launch_bounds__(kBlock) void kb_h_flush(const BSys* __restrict__ sys, int pred) {
    BATCH_SITE_PROLOGUE(pred)
    if (i >= sh) return;
    const double inv = 1.0 / S.st->scale;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const long k = c * sh + i;
        const double2 xx = S.x[k];
        const float2 ff = S.outf[k];
        S.x[k] = make_double2(xx.x + inv * (double)ff.x, xx.y + inv * (double)ff.y);
        S.outf[k] = make_float2(0.f, 0.f);
    }
}

// half copies of one gauge field: rows 0,1 of each link array and its phases as (cos, sin)
struct HLinks {
    const double2* u[8];
    const double* ph[8];
    h8* ula[8];
    h6* ulb[8];
    int sizeh;
};
// *flag = 1 if some phase of the field is not +-1 after rounding to FP16, or
// some row 0 entry 0 real part exceeds 1 in magnitude (every thread that finds
// one stores the same value)
__global__ __launch_bounds__(kBlock) void kb_h_phase_check(HLinks a, int* flag) {
    const long gid = (long)blockIdx.x * kBlock + threadIdx.x;
    if (gid >= 8L * a.sizeh) return;
    const int l = (int)(gid / a.sizeh), i = (int)(gid - (long)l * a.sizeh);
    double sn, cs;
    phase_sincos(a.ph[l][i], &sn, &cs);
    const h2 p = f2h((float)cs, (float)sn);
    const float c = (float)p.x, sv = (float)p.y;
    const float e = (float)(_Float16)(float)a.u[l][i].x;
    if (!(sv == 0.f && (c == 1.f || c == -1.f)) || !(fabsf(e) <= 1.f)) *flag = 1;
}

// FP16 copy of the links: rows 0 and 1 of the FP64 su3_soa planes and the
// phase, as h8 + h6 records (phase as (cos, sin)) or, sgn, h8 + h4 records
// (phase +-1 in bit 14 of row 0 entry 0's real part)
__global__ __launch_bounds__(kBlock) void kb_h_links(HLinks a, int sgn) {
    const long gid = (long)blockIdx.x * kBlock + threadIdx.x;
    if (gid >= 8L * a.sizeh) return;
    const int l = (int)(gid / a.sizeh), i = (int)(gid - (long)l * a.sizeh);
    h2 e[6];
#pragma unroll
    for (int p = 0; p < 6; p++) {
        const double2 v = a.u[l][(long)p * a.sizeh + i];
        e[p] = f2h((float)v.x, (float)v.y);
    }
    double sn, cs;
    phase_sincos(a.ph[l][i], &sn, &cs);
    const h2 ph = f2h((float)cs, (float)sn);
    h8 la;
    la.v[0] = e[0]; la.v[1] = e[1]; la.v[2] = e[2]; la.v[3] = e[3];
    if (sgn) {
        unsigned short bits = __builtin_bit_cast(unsigned short, la.v[0].x);
        bits = (unsigned short)(bits | ((float)ph.x < 0.f ? 0x4000u : 0u));
        la.v[0].x = __builtin_bit_cast(_Float16, bits);
        h4 lc;
        lc.v[0] = e[4]; lc.v[1] = e[5];
        stg(a.ula[l] + i, la);
        stg((h4*)a.ulb[l] + i, lc);
    } else {
        h6 lb;
        lb.v[0] = e[4]; lb.v[1] = e[5]; lb.v[2] = ph;
        stg(a.ula[l] + i, la);
        stg(a.ulb[l] + i, lb);
    }
}

// one block per system: sum its partials in tile order, apply finalizer F
template <int F>
__global__ __launch_bounds__(kBlock) void kb_finalize(const BSys* __restrict__ sys, int pred, int tile_size, int odd) {
    const BSys& S = sys[blockIdx.x];
    const SolverCtl c = bctl(S, pred);
    const int nparts = (S.g.sizeh + tile_size - 1) / tile_size;
    // the partial loads are issued before the predicate's state load (both
    // in bounds), so the two memory latencies overlap
    double a = 0.0;
#pragma unroll 4
    for (int i = threadIdx.x; i < nparts; i += kBlock) a += S.partials[2 * i + odd];
    if (!pred_pass(c)) return;
    __shared__ double lds[kBlock / 64];
    const double t = block_sum<kBlock>(a, lds);
    if (threadIdx.x == 0) finalize<F>(S.st, t, 0.0);
}

// ------------------------------------------------------------------ host side
// Dslash tile of the batched kernels: threads per workgroup and order (as in
// kb_dslash); the inner solver's hops (FP32 or FP16) and the FP64 hops of
// the reliable updates are tuned separately
struct Tile {
    int blk = 256;
    int order = 1;   // bits 0-2 as in kb_dslash; bit 3: FP64 hops by chunks (kb_dslash_mr)
    float us = 0.f;
};
constexpr int kMrSystems = 8;  // systems per chunk of kb_dslash_mr
struct Tiles {
    Tile inner, f64;
};

static std::map<std::string, Tiles> g_tuned;  // batch signature -> chosen Dslash tiles
static std::string g_last_report;

struct Batch {
    int n = 0;
    int max_tiles = 0;   // tiles of the largest system at kBlock threads
    int max_tiles64 = 0; // same at 64 threads (partials capacity)
    BSys* dsys = nullptr;
    SolverState* st = nullptr;
    SolverState* st_host = nullptr;
    SolverState* st_snap[2] = {nullptr, nullptr};  // pinned snapshots taken behind queued replays
    hipEvent_t snap_ev[2] = {nullptr, nullptr};
    void* arena = nullptr;
    size_t arena_bytes = 0;
    size_t st_cap = 0;
    int* chunks = nullptr;  // kMrSystems system indices per chunk (-1 padded), in the arena
    int nchunks = 0;
    int shared = 0;         // some chunk holds more than one system
};
static Batch g_batch;
static bool g_sgn = false;  // FP16 links of this call carry their +-1 phase as a sign bit

// The instantiated graph of the last call, reused while everything its kernel
// arguments depend on is unchanged (the per-system data live behind dsys and
// are uploaded by every call; the scalars live in the device states).
struct GraphKey {
    int nsys = -1, max_sizeh = 0, blk = 0, order = 0, blk64 = 0, order64 = 0, half = 0, sgn = 0, iters = 0;
    int nchunks = 0;
    const void* dsys = nullptr;
    const void* chunks = nullptr;
    bool operator==(const GraphKey& o) const {
        return nsys == o.nsys && max_sizeh == o.max_sizeh && blk == o.blk && order == o.order && blk64 == o.blk64 &&
               order64 == o.order64 && half == o.half && sgn == o.sgn && iters == o.iters && dsys == o.dsys &&
               nchunks == o.nchunks && chunks == o.chunks;
    }
};
static GraphKey g_graph_key;
static hipGraph_t g_graph = nullptr;
static hipGraphExec_t g_graph_exec = nullptr;
static void drop_graph() {
    if (g_graph_exec) (void)hipGraphExecDestroy(g_graph_exec);
    if (g_graph) (void)hipGraphDestroy(g_graph);
    g_graph_exec = nullptr;
    g_graph = nullptr;
    g_graph_key = GraphKey{};
}

static hipStream_t S() { return ctx().stream; }

template <int ROLE>
static void launch_role(const Tile& t, int nsys, int max_sizeh, int pred) {
    const int tiles = (max_sizeh + t.blk - 1) / t.blk;
    if constexpr (ROLE >= kRoleH16Doe) {
        if (t.order & 8) {
            const dim3 grid(tiles, g_batch.nchunks);
            const int o = t.order & 7;
            const int* ch = g_batch.chunks;
            if (t.blk == 64)
                hipLaunchKernelGGL((kb_dslash_h_mr<ROLE, 64, kMrSystems>), grid, dim3(64), 0, S(), g_batch.dsys, ch, pred,
                                   o);
            else if (t.blk == 128)
                hipLaunchKernelGGL((kb_dslash_h_mr<ROLE, 128, kMrSystems>), grid, dim3(128), 0, S(), g_batch.dsys, ch,
                                   pred, o);
            else
                hipLaunchKernelGGL((kb_dslash_h_mr<ROLE, 256, kMrSystems>), grid, dim3(256), 0, S(), g_batch.dsys, ch,
                                   pred, o);
            return;
        }
    }
    if constexpr (ROLE <= kRoleF64DeoResid) {
        if (t.order & 8) {
            const dim3 grid(tiles, g_batch.nchunks);
            const int o = t.order & 7;
            const int* ch = g_batch.chunks;
            if (t.blk == 64)
                hipLaunchKernelGGL((kb_dslash_mr<ROLE, 64, kMrSystems>), grid, dim3(64), 0, S(), g_batch.dsys, ch, pred, o);
            else if (t.blk == 128)
                hipLaunchKernelGGL((kb_dslash_mr<ROLE, 128, kMrSystems>), grid, dim3(128), 0, S(), g_batch.dsys, ch, pred,
                                   o);
            else
                hipLaunchKernelGGL((kb_dslash_mr<ROLE, 256, kMrSystems>), grid, dim3(256), 0, S(), g_batch.dsys, ch, pred,
                                   o);
            return;
        }
    }
    if (t.blk == 64)
        hipLaunchKernelGGL((kb_dslash<ROLE, 64>), dim3(tiles, nsys), dim3(64), 0, S(), g_batch.dsys, pred, t.order);
    else if (t.blk == 128)
        hipLaunchKernelGGL((kb_dslash<ROLE, 128>), dim3(tiles, nsys), dim3(128), 0, S(), g_batch.dsys, pred, t.order);
    else
        hipLaunchKernelGGL((kb_dslash<ROLE, 256>), dim3(tiles, nsys), dim3(256), 0, S(), g_batch.dsys, pred, t.order);
}

// odd = 1: partials stored in the odd slots (kb_h_xr)
template <int F>
static void fin(int nsys, int pred, int tile_size, int odd = 0) {
    hipLaunchKernelGGL((kb_finalize<F>), dim3(nsys), dim3(kBlock), 0, S(), g_batch.dsys, pred, tile_size, odd);
}

static bool pull_states(int n) {
    if (!hip_ok(hipMemcpyAsync(g_batch.st_host, g_batch.st, sizeof(SolverState) * n, hipMemcpyDeviceToHost, S()),
                "state download"))
        return false;
    return hip_ok(hipStreamSynchronize(S()), "stream synchronize");
}
static bool push_states(int n) {
    if (!hip_ok(hipMemcpyAsync(g_batch.st, g_batch.st_host, sizeof(SolverState) * n, hipMemcpyHostToDevice, S()),
                "state upload"))
        return false;
    return hip_ok(hipStreamSynchronize(S()), "stream synchronize");
}

// lattice extents of every system and which systems share a gauge field
// (the index of the first system with the same FP64 links)
static std::string signature(const std::vector<BSys>& v) {
    std::string s;
    for (size_t i = 0; i < v.size(); i++) {
        size_t first = i;
        for (size_t j = 0; j < i; j++)
            if (v[j].u64[0] == v[i].u64[0]) {
                first = j;
                break;
            }
        char buf[64];
        snprintf(buf, sizeof buf, "%dx%dx%dx%d@%zu;", v[i].g.nd0, v[i].g.nd1, v[i].g.nd2, v[i].g.nd3, first);
        s += buf;
    }
    return s;
}

// Time each candidate tile (64, 128 or 256 threads; t-fastest or contiguous
// tiles; XCD-contiguous or hardware order; systems one after the other or
// interleaved) on the batch's inner D_oe (FP32,
// or FP16 for the FP16 inner solver) and on its FP64 D_oe, three launches
// each after one warm-up launch; keep the fastest of each.
static Tile g_forced_tile{0, 0, 0.f};  // blk 0: autotune

static Tiles autotune(const std::string& sig, int nsys, int max_sizeh, bool half) {
    if (g_forced_tile.blk) {
        Tiles t;
        t.inner = t.f64 = g_forced_tile;
        return t;
    }
    auto it = g_tuned.find(sig);
    if (it != g_tuned.end()) return it->second;
    hipEvent_t e0, e1;
    Tiles best;
    best.inner.us = best.f64.us = 1e30f;
    std::string rep = "batch " + sig + ":";
    if (!hip_ok(hipEventCreate(&e0), "event") || !hip_ok(hipEventCreate(&e1), "event")) return Tiles{};
    static const char* const kOrder[16] = {
        "contiguous", "t-fastest", "contiguous hw-order", "t-fastest hw-order",
        "contiguous interleaved", "t-fastest interleaved", "contiguous hw-order interleaved",
        "t-fastest hw-order interleaved", "chunks contiguous", "chunks t-fastest", "chunks contiguous hw-order",
        "chunks t-fastest hw-order", "chunks contiguous interleaved", "chunks t-fastest interleaved",
        "chunks contiguous hw-order interleaved", "chunks t-fastest hw-order interleaved"};
    for (int fam = 0; fam < 2; fam++) {
        rep += fam == 0 ? " inner D_oe:" : " fp64 D_oe:";
        Tile& b = fam == 0 ? best.inner : best.f64;
        const int norders = g_batch.shared ? 16 : 8;  // chunked hops only where systems share a field
        for (int blk : {256, 128, 64})
            for (int order = 0; order < norders; order++) {
                Tile c;
                c.blk = blk;
                c.order = order;
                auto one = [&]() {
                    if (fam == 1) launch_role<kRoleF64Doe>(c, nsys, max_sizeh, kAlways);
                    else if (half && g_sgn) launch_role<kRoleS16Doe>(c, nsys, max_sizeh, kAlways);
                    else if (half) launch_role<kRoleH16Doe>(c, nsys, max_sizeh, kAlways);
                    else launch_role<kRoleF32Doe>(c, nsys, max_sizeh, kAlways);
                };
                one();
                (void)hipEventRecord(e0, S());
                for (int r = 0; r < 3; r++) one();
                (void)hipEventRecord(e1, S());
                (void)hipEventSynchronize(e1);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                c.us = 1000.f * ms / 3;
                char buf[96];
                snprintf(buf, sizeof buf, " [%d thr, %s] %.1f us", c.blk, kOrder[c.order], c.us);
                rep += buf;
                if (c.us < b.us) b = c;
            }
        char buf[96];
        snprintf(buf, sizeof buf, " -> %d thr %s;", b.blk, kOrder[b.order]);
        rep += buf;
    }
    rep += "\n";
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g_last_report += rep;
    if (verbosity_lv > 1) printf("%s", rep.c_str());
    g_tuned[sig] = best;
    return best;
}

}  // namespace ostp

using namespace ostp;

namespace ostp {

static int batched_solve(int nsys, const ostp_batch_system* sys, double res, int max_cg, double mixed_delta,
                         int iters_per_graph, int* cg_return, int* ru_return, double* true_res, bool half) {
    const char* nm = half ? "ostp_batched_half_solve" : "ostp_batched_mixed_solve";
    if (!require_init(nm)) return 0;
    if (nsys <= 0 || !sys || iters_per_graph < 0) {
        fail("%s: invalid arguments", nm);
        return 0;
    }
    // ---- descriptors
    std::vector<BSys> hs(nsys);
    std::vector<long> sizeh(nsys);
    size_t arena = 0;
    int max_sizeh = 0, max_t64 = 0;
    for (int k = 0; k < nsys; k++) {
        const ostp_batch_system& q = sys[k];
        const int* nd = q.nd;
        if (nd[0] <= 0 || nd[1] <= 0 || nd[2] <= 0 || nd[3] <= 0 || (nd[0] & 1)) {
            fail("%s: system %d has invalid extents", nm, k);
            return 0;
        }
        sizeh[k] = (long)nd[0] * nd[1] * nd[2] * nd[3] / 2;
        if (sizeh[k] > (1L << 29)) {
            fail("%s: system %d too large", nm, k);
            return 0;
        }
        max_sizeh = std::max(max_sizeh, (int)sizeh[k]);
        max_t64 = std::max(max_t64, (int)((sizeh[k] + 63) / 64));
        arena += (size_t)(3 * 48 + 5 * 24) * sizeh[k] + 256 * 8;  // vectors (+ alignment slack)
        if (half) arena += (size_t)(4 * 12 + 8 * 28) * sizeh[k] + 256 * 21;  // half vectors + a half link copy + flag
    }
    const size_t part_per = (size_t)2 * max_t64;
    arena += (sizeof(double) * part_per + 256) * nsys;
    arena += sizeof(int) * kMrSystems * (size_t)nsys + 256;  // chunk table
    Batch& B = g_batch;
    if (arena > B.arena_bytes) {
        if (B.arena) (void)hipFree(B.arena);
        B.arena = nullptr;
        B.arena_bytes = 0;
        if (!hip_ok(hipMalloc(&B.arena, arena), "batch arena")) return 0;
        B.arena_bytes = arena;
    }
    if ((size_t)nsys > B.st_cap) {
        if (B.st) (void)hipFree(B.st);
        if (B.st_host) (void)hipHostFree(B.st_host);
        for (int j = 0; j < 2; j++)
            if (B.st_snap[j]) (void)hipHostFree(B.st_snap[j]);
        if (B.dsys) (void)hipFree(B.dsys);
        B.st = nullptr;
        B.st_host = nullptr;
        B.st_snap[0] = B.st_snap[1] = nullptr;
        B.dsys = nullptr;
        B.st_cap = 0;
        if (!hip_ok(hipMalloc(&B.st, sizeof(SolverState) * nsys), "batch states") ||
            !hip_ok(hipHostMalloc(&B.st_host, sizeof(SolverState) * nsys, hipHostMallocDefault), "batch states") ||
            !hip_ok(hipHostMalloc(&B.st_snap[0], sizeof(SolverState) * nsys, hipHostMallocDefault), "batch states") ||
            !hip_ok(hipHostMalloc(&B.st_snap[1], sizeof(SolverState) * nsys, hipHostMallocDefault), "batch states") ||
            !hip_ok(hipMalloc(&B.dsys, sizeof(BSys) * nsys), "batch descriptors"))
            return 0;
        B.st_cap = nsys;
    }
    char* cur = (char*)B.arena;
    auto take = [&](size_t bytes) {
        void* p = cur;
        cur += (bytes + 255) & ~(size_t)255;
        return p;
    };
    for (int k = 0; k < nsys; k++) {
        const ostp_batch_system& q = sys[k];
        BSys& b = hs[k];
        const long sh = sizeh[k];
        const size_t su3 = (size_t)(144 + (q.par_temp ? 8 : 0)) * sh, su3f = (size_t)(72 + (q.par_temp ? 4 : 0)) * sh;
        const char* u = (const char*)resolve(q.u, 8 * su3, nm);
        const char* uf = (const char*)resolve(q.u_f, 8 * su3f, nm);
        const char* ph = (const char*)resolve(q.phases, 8 * 8 * sh, nm);
        const char* phf = (const char*)resolve(q.phases_f, 8 * 4 * sh, nm);
        b.b = (const double2*)resolve(q.in, 48 * sh, nm);
        b.x = (double2*)resolve(q.out, 48 * sh, nm);
        if (!u || !uf || !ph || !phf || !b.b || !b.x) return 0;
        for (int l = 0; l < 8; l++) {
            b.u64[l] = (const double2*)(u + l * su3);
            b.u32[l] = (const float2*)(uf + l * su3f);
            b.ph64[l] = (const double*)(ph + (size_t)l * 8 * sh);
            b.ph32[l] = (const float*)(phf + (size_t)l * 4 * sh);
        }
        b.r64 = (double2*)take(48 * sh);
        b.h64 = (double2*)take(48 * sh);
        b.s64 = (double2*)take(48 * sh);
        b.rf = (float2*)take(24 * sh);
        b.hf = (float2*)take(24 * sh);
        b.sf = (float2*)take(24 * sh);
        b.pf = (float2*)take(24 * sh);
        b.outf = (float2*)take(24 * sh);
        b.g.nd0 = q.nd[0];
        b.g.nd0h = q.nd[0] / 2;
        b.g.nd1 = q.nd[1];
        b.g.nd2 = q.nd[2];
        b.g.nd3 = q.nd[3];
        b.g.sizeh = (int)sh;
        b.m2 = q.mass * q.mass;
        b.m2f = q.mass * q.mass;  // inverter_mixedp.c passes shift 0
        b.st = B.st + k;
        b.partials = nullptr;
    }
    for (int k = 0; k < nsys; k++) hs[k].partials = (double*)take(sizeof(double) * part_per);
    {
        // chunks of up to kMrSystems systems on one gauge field (same FP64 and
        // FP32 links and link strides, phases and extents), in system order
        std::vector<int> tab;
        std::vector<char> done(nsys, 0);
        int shared = 0;
        for (int k = 0; k < nsys; k++) {
            if (done[k]) continue;
            std::vector<int> grp;
            for (int j = k; j < nsys; j++)
                if (!done[j] && hs[j].u64[0] == hs[k].u64[0] && hs[j].u64[1] == hs[k].u64[1] &&
                    hs[j].ph64[0] == hs[k].ph64[0] && hs[j].u32[0] == hs[k].u32[0] &&
                    hs[j].u32[1] == hs[k].u32[1] && hs[j].ph32[0] == hs[k].ph32[0] &&
                    hs[j].g.nd0 == hs[k].g.nd0 && hs[j].g.nd1 == hs[k].g.nd1 && hs[j].g.nd2 == hs[k].g.nd2 &&
                    hs[j].g.nd3 == hs[k].g.nd3) {
                    grp.push_back(j);
                    done[j] = 1;
                }
            for (size_t a = 0; a < grp.size(); a += kMrSystems) {
                for (int r = 0; r < kMrSystems; r++)
                    tab.push_back(a + r < grp.size() ? grp[a + r] : -1);
                if (grp.size() - a > 1) shared = 1;
            }
        }
        B.nchunks = (int)(tab.size() / kMrSystems);
        B.shared = shared;
        B.chunks = (int*)take(sizeof(int) * tab.size());
        if (!hip_ok(hipMemcpyAsync(B.chunks, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice, S()),
                    "chunk table"))
            return 0;
    }
    if (half) {
        // half vectors, and one half copy of each distinct gauge field (systems sharing links share it)
        std::map<std::pair<const void*, const void*>, int> seen;
        std::vector<HLinks> fields;
        for (int k = 0; k < nsys; k++) {
            BSys& b = hs[k];
            const long sh = sizeh[k];
            b.rh = (h6*)take(12 * sh);
            b.hh = (h6*)take(12 * sh);
            b.sh = (h6*)take(12 * sh);
            b.pph = (h6*)take(12 * sh);
            const auto key = std::make_pair((const void*)b.u64[0], (const void*)b.ph64[0]);
            auto it = seen.find(key);
            if (it != seen.end()) {
                for (int l = 0; l < 8; l++) {
                    b.ula[l] = hs[it->second].ula[l];
                    b.ulb[l] = hs[it->second].ulb[l];
                    b.ulc[l] = hs[it->second].ulc[l];
                }
                continue;
            }
            seen[key] = k;
            HLinks a;
            for (int l = 0; l < 8; l++) {
                a.u[l] = b.u64[l];
                a.ph[l] = b.ph64[l];
                a.ula[l] = (h8*)take((size_t)16 * sh);
                a.ulb[l] = (h6*)take((size_t)12 * sh);  // room for either second record
                b.ula[l] = a.ula[l];
                b.ulb[l] = a.ulb[l];
                b.ulc[l] = (const h4*)a.ulb[l];
            }
            a.sizeh = (int)sh;
            fields.push_back(a);
        }
        // the 24 B sign-carrying links if every phase of every field is +-1
        int* flag = (int*)take(256);
        int h_flag = 0;
        if (!hip_ok(hipMemsetAsync(flag, 0, sizeof(int), S()), "phase flag")) return 0;
        for (const HLinks& a : fields)
            hipLaunchKernelGGL(kb_h_phase_check, dim3(blocks_for(8L * a.sizeh)), dim3(kBlock), 0, S(), a, flag);
        if (!hip_ok(hipMemcpyAsync(&h_flag, flag, sizeof(int), hipMemcpyDeviceToHost, S()), "phase flag") ||
            !hip_ok(hipStreamSynchronize(S()), "phase flag"))
            return 0;
        g_sgn = (h_flag == 0);
        for (const HLinks& a : fields)
            hipLaunchKernelGGL(kb_h_links, dim3(blocks_for(8L * a.sizeh)), dim3(kBlock), 0, S(), a, (int)g_sgn);
        if (!hip_ok(hipGetLastError(), "half links")) return 0;
    }
    if (!hip_ok(hipMemcpyAsync(B.dsys, hs.data(), sizeof(BSys) * nsys, hipMemcpyHostToDevice, S()), "descriptors"))
        return 0;
    const int nb_site = (max_sizeh + kBlock - 1) / kBlock;
    const dim3 gsite(nb_site, nsys);
    const dim3 gxr((max_sizeh + kBlock * kXrSites - 1) / (kBlock * kXrSites), nsys);

    // ---- source norms, states
    hipLaunchKernelGGL(kb_resid_norm, gsite, dim3(kBlock), 0, S(), B.dsys, 0);
    fin<kFinStore>(nsys, kAlways, kBlock);
    if (!hip_ok(hipGetLastError(), nm) || !pull_states(nsys)) return 0;
    for (int k = 0; k < nsys; k++) {
        SolverState& s = B.st_host[k];
        const double sn = s.sum;
        std::memset(&s, 0, sizeof s);
        s.res = res;
        s.safety = 0.9;
        s.max_cg = max_cg;
        s.mixed_delta = mixed_delta;
        s.source_norm = sn;
        s.nentries = 3.0 * (double)sizeh[k];
        s.status = kRunning;
    }
    if (!push_states(nsys)) return 0;
    const std::string sig = signature(hs) + (half ? (g_sgn ? " fp16 sign-phase" : " fp16") : " fp32");
    const Tiles tt = autotune(sig, nsys, max_sizeh, half);
    const Tile& t = tt.inner;  // the inner solver's hops
    const Tile& t64 = tt.f64;  // the FP64 hops

    // r = b - M x (FP64); inner r = p = r, out = 0, delta = |r_inner|^2
    launch_role<kRoleF64Doe>(t64, nsys, max_sizeh, kAlways);
    if (half) {
        launch_role<kRoleF64DeoResid>(t64, nsys, max_sizeh, kAlways);
        fin<kFinHScale>(nsys, kAlways, t64.blk);
        hipLaunchKernelGGL(kb_h_rescale, gsite, dim3(kBlock), 0, S(), B.dsys, (int)kAlways, 1);
        fin<kFinHInit>(nsys, kAlways, kBlock);
    } else {
        launch_role<kRoleF64Deo>(t64, nsys, max_sizeh, kAlways);
        hipLaunchKernelGGL(kb_mx_init, gsite, dim3(kBlock), 0, S(), B.dsys);
        fin<kFinCgRestart>(nsys, kAlways, kBlock);
    }
    if (!hip_ok(hipGetLastError(), nm)) return 0;

    // ---- one inner iteration of every system (predicated on each system's state)
    auto enqueue_iteration = [&]() {
        if (half) {
            if (g_sgn) {
                launch_role<kRoleS16Doe>(t, nsys, max_sizeh, kIfGoing);
                launch_role<kRoleS16DeoDot>(t, nsys, max_sizeh, kIfGoing);
            } else {
                launch_role<kRoleH16Doe>(t, nsys, max_sizeh, kIfGoing);
                launch_role<kRoleH16DeoDot>(t, nsys, max_sizeh, kIfGoing);
            }
            hipLaunchKernelGGL(kb_h_xr, gxr, dim3(kBlock), 0, S(), B.dsys, t.blk);
            // p = p gammag + r happens inside the next iteration's two hops
        } else {
            launch_role<kRoleF32Doe>(t, nsys, max_sizeh, kIfRunning);
            launch_role<kRoleF32DeoDot>(t, nsys, max_sizeh, kIfRunning);
            fin<kFinMxAlpha>(nsys, kIfRunning, t.blk);
            hipLaunchKernelGGL(kb_mx_xr, gsite, dim3(kBlock), 0, S(), B.dsys);
            fin<kFinMxLambda>(nsys, kIfRunningNoRU, kBlock);
            launch_role<kRoleF64Doe>(t64, nsys, max_sizeh, kIfRunningRU);
            launch_role<kRoleF64Deo>(t64, nsys, max_sizeh, kIfRunningRU);
            hipLaunchKernelGGL(kb_mx_resid, gsite, dim3(kBlock), 0, S(), B.dsys);
            fin<kFinMxLambda>(nsys, kIfRunningRU, kBlock);
            hipLaunchKernelGGL(kb_cg_p, gsite, dim3(kBlock), 0, S(), B.dsys);
        }
    };
    // ---- iters_per_graph iterations captured once into a HIP graph (0: the same
    // sequence launched directly, 16 iterations per round, for tools that do not
    // follow stream capture)
    // FP16 inner solver: the reliable update (due, or asked for by the inner
    // exit test) runs once after each block of queued iterations; the
    // iterations in between continue the inner recurrence
    auto enqueue_reliable_update = [&]() {
        if (!half) return;
        hipLaunchKernelGGL(kb_h_flush, gsite, dim3(kBlock), 0, S(), B.dsys, (int)kIfRunningRU);
        launch_role<kRoleF64Doe>(t64, nsys, max_sizeh, kIfRunningRU);
        launch_role<kRoleF64DeoResid>(t64, nsys, max_sizeh, kIfRunningRU);
        fin<kFinHScale>(nsys, kIfRunningRU, t64.blk);
        hipLaunchKernelGGL(kb_h_rescale, gsite, dim3(kBlock), 0, S(), B.dsys, (int)kIfRunningRU, 0);
        fin<kFinHRu>(nsys, kIfRunningRU, kBlock);
    };
    hipGraphExec_t exec = nullptr;
    if (iters_per_graph > 0) {
        GraphKey key;
        key.nsys = nsys;
        key.max_sizeh = max_sizeh;
        key.blk = t.blk;
        key.order = t.order;
        key.blk64 = t64.blk;
        key.order64 = t64.order;
        key.half = half ? 1 : 0;
        key.sgn = g_sgn ? 1 : 0;
        key.iters = iters_per_graph;
        key.dsys = B.dsys;
        key.nchunks = B.nchunks;
        key.chunks = B.chunks;
        if (!(g_graph_exec && key == g_graph_key)) {
            // the stream is idle here (the states were pulled), so no replay of the old graph runs
            drop_graph();
            hipGraph_t graph = nullptr;
            if (!hip_ok(hipStreamBeginCapture(S(), hipStreamCaptureModeThreadLocal), "begin capture")) return 0;
            // FP16 inner solver: a reliable-update tail after every kRuBlock
            // iterations and at the end, so a long graph (a whole solve in
            // one launch) keeps the FP16 recurrence near its stored scale
            for (int j = 0; j < iters_per_graph; j++) {
                enqueue_iteration();
                if ((j + 1) % kRuBlock == 0 && j + 1 < iters_per_graph) enqueue_reliable_update();
            }
            enqueue_reliable_update();
            const hipError_t ce = hipStreamEndCapture(S(), &graph);
            if (!hip_ok(ce, "end capture")) return 0;
            if (!hip_ok(hipGraphInstantiate(&g_graph_exec, graph, nullptr, nullptr, 0), "graph instantiate")) {
                (void)hipGraphDestroy(graph);
                g_graph_exec = nullptr;
                return 0;
            }
            g_graph = graph;
            g_graph_key = key;
        }
        exec = g_graph_exec;
    }
    // Replays are pipelined: replay k+1 is queued (with a state snapshot
    // behind it) before the host waits for the snapshot taken after replay
    // k, so the GPU never idles on the host's convergence check.  The replay
    // queued past the end runs as predicated no-ops.
    for (int j = 0; j < 2; j++)
        if (!B.snap_ev[j] && !hip_ok(hipEventCreateWithFlags(&B.snap_ev[j], hipEventDisableTiming), "event"))
            return 0;
    auto replay = [&](int slot) -> bool {
        if (exec) {
            if (!hip_ok(hipGraphLaunch(exec, S()), "graph launch")) return false;
        } else {
            for (int j = 0; j < 16; j++) enqueue_iteration();
            enqueue_reliable_update();
            if (!hip_ok(hipGetLastError(), nm)) return false;
        }
        return hip_ok(hipMemcpyAsync(B.st_snap[slot], B.st, sizeof(SolverState) * nsys, hipMemcpyDeviceToHost, S()),
                      "state snapshot") &&
               hip_ok(hipEventRecord(B.snap_ev[slot], S()), "event");
    };
    bool ok = true;
    long last_total = -1;
    bool any = false;
    for (int k = 0; k < nsys; k++) any = any || B.st_host[k].status == kRunning;
    if (any) {
        int cur = 0;
        ok = replay(cur);
        while (ok) {
            if (!replay(cur ^ 1) || !hip_ok(hipEventSynchronize(B.snap_ev[cur]), "snapshot wait")) {
                ok = false;
                break;
            }
            long total = 0;
            bool running = false;
            for (int k = 0; k < nsys; k++) {
                total += B.st_snap[cur][k].cg;
                running = running || B.st_snap[cur][k].status == kRunning;
            }
            if (!running) break;
            if (total == last_total) {
                ok = fail("%s: solver states did not advance", nm);
                break;
            }
            last_total = total;
            cur ^= 1;
        }
    }
    // the replay queued past the end may still run: drain the stream before
    // anything it reads is released or rewritten
    if (!hip_ok(hipStreamSynchronize(S()), "drain replays")) ok = false;
    if (!ok) {
        drop_graph();
        return 0;
    }

    // ---- x += out_f; FP64 true residual of every system.  The FP16 inner
    // solver ends every system on a reliable update (or at its start, when
    // that is already within the tolerance), whose FP64 residual is that of
    // the final x: it is recomputed only if some system stopped elsewhere
    // (at max_cg inside an inner iteration).
    bool fresh = false;
    if (half) {
        if (!pull_states(nsys)) return 0;
        fresh = true;
        for (int k = 0; k < nsys; k++) fresh = fresh && B.st_host[k].resid_ok == 1;
    }
    if (!fresh) {
        if (half)
            hipLaunchKernelGGL(kb_h_flush, gsite, dim3(kBlock), 0, S(), B.dsys, (int)kAlways);
        else
            hipLaunchKernelGGL(kb_add_f_to_d, gsite, dim3(kBlock), 0, S(), B.dsys);
        launch_role<kRoleF64Doe>(t64, nsys, max_sizeh, kAlways);
        launch_role<kRoleF64Deo>(t64, nsys, max_sizeh, kAlways);
        hipLaunchKernelGGL(kb_resid_norm, gsite, dim3(kBlock), 0, S(), B.dsys, 1);
        fin<kFinStore>(nsys, kAlways, kBlock);
        if (!hip_ok(hipGetLastError(), nm) || !pull_states(nsys)) return 0;
    }
    int converged = 0;
    for (int k = 0; k < nsys; k++) {
        const SolverState& s = B.st_host[k];
        const double tr = std::sqrt(s.sum / s.source_norm);
        if (cg_return) cg_return[k] = s.cg;
        if (ru_return) ru_return[k] = s.ru_count;
        if (true_res) true_res[k] = tr;
        if (tr <= res) converged++;
        else if (verbosity_lv > 0)
            printf("WARNING: batched system %d: true residual %e above %e after %d iterations\n", k, tr, res, s.cg);
    }
    return converged;
}

int batched_solve_internal(int nsys, const ostp_batch_system* sys, double res, int max_cg, double mixed_delta,
                           int iters_per_graph, int* cg_return, int* ru_return, double* true_res, bool half) {
    return batched_solve(nsys, sys, res, max_cg, mixed_delta, iters_per_graph, cg_return, ru_return, true_res, half);
}

}  // namespace ostp

extern "C" {

int ostp_batched_mixed_solve(int nsys, const ostp_batch_system* sys, double res, int max_cg, double mixed_delta,
                             int iters_per_graph, int* cg_return, int* ru_return, double* true_res) {
    return batched_solve(nsys, sys, res, max_cg, mixed_delta, iters_per_graph, cg_return, ru_return, true_res, false);
}

int ostp_batched_half_solve(int nsys, const ostp_batch_system* sys, double res, int max_cg, double mixed_delta,
                            int iters_per_graph, int* cg_return, int* ru_return, double* true_res) {
    return batched_solve(nsys, sys, res, max_cg, mixed_delta, iters_per_graph, cg_return, ru_return, true_res, true);
}

int ostp_set_batch_tile(int blk, int order) {
    if (blk != 0 && ((blk != 64 && blk != 128 && blk != 256) || order < 0 || order > 15)) {
        fail("ostp_set_batch_tile: workgroup %d / order %d (64, 128 or 256 and 0..15, or 0 for the autotuner)", blk,
             order);
        return -1;
    }
    const int prev = g_forced_tile.blk ? g_forced_tile.blk * 16 + g_forced_tile.order : 0;
    g_forced_tile.blk = blk;
    g_forced_tile.order = blk ? order : 0;
    return prev;
}

int ostp_autotune_report(char* buf, int len) {
    const int n = (int)g_last_report.size();
    if (buf && len > 0) {
        const int m = std::min(n, len - 1);
        std::memcpy(buf, g_last_report.data(), m);
        buf[m] = 0;
    }
    return n;
}

}  // extern "C"
