// Learned gauge-equivariant preconditioner for the even-odd staggered system
// A = m^2 - D_eo D_oe, and the flexible CG that uses it.
//
// Network (2 parallel-transport layers, C = 12 channels, complex weights
// w[c][nu], real w0):
//   layer 1 (even -> odd)  h_c = sum_nu w[c][nu] T_nu r
//   layer 2 (odd -> even)  z   = w0 r + sum_nu T_nu^dag g_nu,  g_nu = sum_c conj(w[c][nu]) h_c
// with T_nu the eight parallel transports of D_oe (fermion_matrix.c:101-157,
// before its 1/2 and signs): N = w0 + sum_c B_c^dag B_c is Hermitian (positive
// for w0 > 0; trained networks have w0 < 0 with N still positive on the
// spectrum), as conjugate-gradient methods need.  Since g = W^dag W t, the
// channel mixing of a site and colour is one Hermitian 8 x 8 product
// g = M t, M = W^dag W, run on the matrix cores as a 32x32x16 BF16 MFMA with
// FP32 accumulation (the real 16 x 16 form of M, 32 columns = 32 colour
// vectors per MFMA; the transported vectors reach the operand layout through
// LDS).  Links are FP16 records (ostp_half.h), transports FP32 arithmetic,
// the channel fields g BF16 records; the inputs and outputs stay FP64.
//
// Training kernels: the loss |e|^2 or <e, A e> (e = N b - x) over (b, x)
// pairs, its gradient via the 8x8 Gram matrix P[nu][mu] = <T_nu a, T_mu b>
// with a = e or A e.  The CPU
// oracle (oracle/ptc.c) restates forward, loss and gradient in FP64.
#include "ostp_dslash.h"
#include "ostp_half.h"

#include <functional>
#include <cmath>
#include <cstring>
#include <vector>

namespace ostp {

constexpr int kC = OSTP_PTC_CHANNELS;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct PtcParams {
    float w0;
    float w[kC][8][2];  // (re, im)
};

// The channel mixing g = W^dag (W t) of every site and colour is g = M t with
// the 8 x 8 Hermitian M = W^dag W; its real 16 x 16 form (rows 2 nu + re/im,
// columns 2 mu + re/im) is the A operand of one 32x32x16 BF16 MFMA per 32
// columns (rows 16..31 zero).
struct PtcMix {
    float m[16][16];
};

// links of the preconditioner: the FP16 records of ostp_half.h (rows 0, 1 and
// the phase), built from the FP32 link copy and FP32 phases by k_ptc_links
struct PtcGeom {
    const h8* la[8];
    const h6* lb[8];  // (cos, sin) phases; with SGN read as h4 (+-1 phases in a sign bit)
    SiteGeom g;
};

struct V3 {
    float2 c[3];
};

// Channel fields g_nu of the odd sites: three complex BF16 numbers per record
// (re in the low half of each dword).  BF16 keeps FP32's exponent range, so
// the fields of a residual that has fallen by 10 orders of magnitude stay
// normal numbers (FP16 would flush them to zero); their 8-bit mantissa
// matches the MFMA operands they come from.
struct __attribute__((aligned(4))) b6 {
    unsigned int v[3];
};
__device__ __forceinline__ unsigned int bf_pack(float x, float y) {
    return (unsigned int)__builtin_bit_cast(unsigned short, (__bf16)x) |
           ((unsigned int)__builtin_bit_cast(unsigned short, (__bf16)y) << 16);
}
__device__ __forceinline__ float2 bf_unpack(unsigned int w) {
    return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}
typedef __attribute__((address_space(1))) const unsigned int gcu32;
typedef __attribute__((address_space(1))) unsigned int gu32;
__device__ __forceinline__ b6 ldg(const b6* p) {
    gcu32* g = (gcu32*)p;
    b6 r;
    r.v[0] = g[0]; r.v[1] = g[1]; r.v[2] = g[2];
    return r;
}
__device__ __forceinline__ void stg(b6* p, const b6& v) {
    gu32* g = (gu32*)p;
    g[0] = v.v[0]; g[1] = v.v[1]; g[2] = v.v[2];
}

__device__ __forceinline__ float2 ldv(const double2* p) {
    const double2 v = *p;
    return make_float2((float)v.x, (float)v.y);
}
__device__ __forceinline__ float2 ldv(const float2* p) { return *p; }
// a vector entry (FP64 or FP32 field) in double precision and back
__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }
template <typename TR> __device__ __forceinline__ TR from_d2(double2 v);
template <> __device__ __forceinline__ double2 from_d2<double2>(double2 v) { return v; }
template <> __device__ __forceinline__ float2 from_d2<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }

// One transport: U e^{i th} v (DAG = false) or U^dag e^{-i th} v (DAG = true)
// with the link of array l at half-site iu
template <bool DAG, bool SGN>
__device__ __forceinline__ V3 xport(const PtcGeom& A, int l, int iu, float2 v0, float2 v1, float2 v2) {
    const h8 la = ldg(A.la[l] + iu);
    float2 a0 = h2f(la.v[0]);
    const float2 a1 = h2f(la.v[1]), a2 = h2f(la.v[2]), b0 = h2f(la.v[3]);
    float2 b1, b2;
    if constexpr (SGN) {
        const h4 lc = ldg((const h4*)A.lb[l] + iu);
        b1 = h2f(lc.v[0]);
        b2 = h2f(lc.v[1]);
        const unsigned short bits = __builtin_bit_cast(unsigned short, la.v[0].x);
        a0.x = (float)__builtin_bit_cast(_Float16, (unsigned short)(bits & 0xbfffu));
        if (bits & 0x4000u) {
            v0 = make_float2(-v0.x, -v0.y);
            v1 = make_float2(-v1.x, -v1.y);
            v2 = make_float2(-v2.x, -v2.y);
        }
    } else {
        const h6 lb = ldg(A.lb[l] + iu);
        b1 = h2f(lb.v[0]);
        b2 = h2f(lb.v[1]);
        const float2 cs = h2f(lb.v[2]);
        const float2 ph = DAG ? make_float2(cs.x, -cs.y) : cs;
        v0 = cmul(v0, ph);
        v1 = cmul(v1, ph);
        v2 = cmul(v2, ph);
    }
    float2 o0 = make_float2(0.f, 0.f), o1 = o0, o2 = o0;
    hop_rows<DAG>(a0, a1, a2, b0, b1, b2, v0, v1, v2, o0, o1, o2);  // DAG subtracts
    V3 o;
    if (DAG) {
        o.c[0] = make_float2(-o0.x, -o0.y);
        o.c[1] = make_float2(-o1.x, -o1.y);
        o.c[2] = make_float2(-o2.x, -o2.y);
    } else {
        o.c[0] = o0;
        o.c[1] = o1;
        o.c[2] = o2;
    }
    return o;
}

template <bool DAG, bool SGN, typename TR>
__device__ __forceinline__ V3 xport(const PtcGeom& A, int l, int iu, const TR* __restrict__ in, int sizeh,
                                    int iv) {
    return xport<DAG, SGN>(A, l, iu, ldv(in + iv), ldv(in + sizeh + iv), ldv(in + 2 * sizeh + iv));
}

// The eight transports T_nu in (in: even FP64 or FP32 field) at odd
// half-site y (the hops of D_oe, fermion_matrix.c:101-157, before its 1/2
// and signs).
template <bool SGN, typename TR>
__device__ __forceinline__ void transports(const PtcGeom& A, const TR* in, int y, V3 t[8]) {
    const Nbrs nb = site_neighbours<1>(A.g, y);
    const int sizeh = A.g.sizeh;
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
        t[mu] = xport<false, SGN>(A, 2 * mu + 1, y, in, sizeh, nb.ip[mu]);
        t[mu + 4] = xport<true, SGN>(A, 2 * mu, nb.im[mu], in, sizeh, nb.im[mu]);
    }
}

// FP16 link records of the preconditioner from the FP32 links and phases
// (sgn: every phase is +-1, carried in a sign bit; flag: set when some phase
// is not, or some row 0 entry 0 exceeds 1 in magnitude).  A phase counts as
// +-1 when its cosine rounds to +-1 in FP16 and |sin| <= 2^-20: the FP32
// argument pi has sin = -8.7e-8, which the sign form drops (far below the
// FP16 resolution of the links, 2^-11).
struct PtcLinkSrc {
    const float2* u[8];
    const float* ph[8];
    h8* la[8];
    h6* lb[8];
    int sizeh;
};
__global__ __launch_bounds__(kBlock) void k_ptc_links_check(PtcLinkSrc a, int* flag) {
    const long gid = (long)blockIdx.x * kBlock + threadIdx.x;
    if (gid >= 8L * a.sizeh) return;
    const int l = (int)(gid / a.sizeh), i = (int)(gid - (long)l * a.sizeh);
    float sn, cs;
    sincos_t(a.ph[l][i], &sn, &cs);
    const float c = (float)(_Float16)cs;
    const float e = (float)(_Float16)a.u[l][i].x;
    if (!(fabsf(sn) <= 0x1p-20f && (c == 1.f || c == -1.f)) || !(fabsf(e) <= 1.f)) *flag = 1;
}
__global__ __launch_bounds__(kBlock) void k_ptc_links(PtcLinkSrc a, int sgn) {
    const long gid = (long)blockIdx.x * kBlock + threadIdx.x;
    if (gid >= 8L * a.sizeh) return;
    const int l = (int)(gid / a.sizeh), i = (int)(gid - (long)l * a.sizeh);
    h2 e[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const float2 v = a.u[l][(long)q * a.sizeh + i];
        e[q] = f2h(v.x, v.y);
    }
    float sn, cs;
    sincos_t(a.ph[l][i], &sn, &cs);
    const h2 ph = f2h(cs, sn);
    h8 la;
    la.v[0] = e[0]; la.v[1] = e[1]; la.v[2] = e[2]; la.v[3] = e[3];
    if (sgn) {
        unsigned short bits = __builtin_bit_cast(unsigned short, la.v[0].x);
        bits = (unsigned short)(bits | ((float)ph.x < 0.f ? 0x4000u : 0u));
        la.v[0].x = __builtin_bit_cast(_Float16, bits);
        h4 lc;
        lc.v[0] = e[4]; lc.v[1] = e[5];
        stg(a.la[l] + i, la);
        stg((h4*)a.lb[l] + i, lc);
    } else {
        h6 lb;
        lb.v[0] = e[4]; lb.v[1] = e[5]; lb.v[2] = ph;
        stg(a.la[l] + i, la);
        stg(a.lb[l] + i, lb);
    }
}

// Layer 1 + channel mixing: g_nu(y) = sum_mu M[nu][mu] T_mu r (y) at odd y,
// stored as one BF16 record (3 colours) per channel nu and half-site.
// One wave = 64 odd sites; MFMA columns n = colour * 64 + site-in-wave.
template <bool SGN, typename TR>
__global__ __launch_bounds__(256) void k_ptc_layer1(PtcGeom A, const TR* __restrict__ r, b6* __restrict__ gout,
                                                    PtcMix M, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    __shared__ bf16x8 lds[4][192 * 2];  // per wave: [n][k], k = 2 mu + re/im (16 bf16 per column)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int base = blockIdx.x * 256 + wave * 64;
    const int y = base + lane;
    const int sizeh = A.g.sizeh;
    V3 t[8];
    if (y < sizeh) {
        transports<SGN>(A, r, y, t);
    } else {
#pragma unroll
        for (int mu = 0; mu < 8; mu++) t[mu].c[0] = t[mu].c[1] = t[mu].c[2] = make_float2(0.f, 0.f);
    }
    bf16x8* T = lds[wave];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int n = c * 64 + lane;
#pragma unroll
        for (int q = 0; q < 2; q++) {  // k = 8q .. 8q+7: mu = 4q .. 4q+3, (re, im) each
            bf16x8 v;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                v[2 * e] = (__bf16)t[4 * q + e].c[c].x;
                v[2 * e + 1] = (__bf16)t[4 * q + e].c[c].y;
            }
            T[n * 2 + q] = v;
        }
    }
    __syncthreads();
    // A operand: row m = lane % 32, k = 8 (lane / 32) + e
    const int m = lane & 31, h = lane >> 5;
    bf16x8 a;
#pragma unroll
    for (int e = 0; e < 8; e++) a[e] = (__bf16)(m < 16 ? M.m[m][8 * h + e] : 0.f);
    // result: lane holds column n = 32 j + m, rows 8 (v / 4) + 4 h + v % 4; rows < 16 are (nu, re/im):
    // v = 0..7 -> nu = 2h + (v / 2) % 2 (v < 4) or 4 + 2h + (v / 2) % 2 (4 <= v < 8)
    float2 acc[2][4][3];  // [site parity p][nu slot][colour]
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const int n = j * 32 + m;
        f32x16 d = {0.f};
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, T[n * 2 + h], d, 0, 0, 0);
        const int p = j & 1, c = j >> 1;  // site = m + 32 p, colour c
#pragma unroll
        for (int q = 0; q < 4; q++) acc[p][q][c] = make_float2(d[2 * q], d[2 * q + 1]);
    }
#pragma unroll
    for (int p = 0; p < 2; p++) {
        const int site = base + m + 32 * p;
        if (site < sizeh) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int nu = (q < 2 ? 0 : 4) + 2 * h + (q & 1);
                b6 rec;
#pragma unroll
                for (int c = 0; c < 3; c++) rec.v[c] = bf_pack(acc[p][q][c].x, acc[p][q][c].y);
                stg(gout + (size_t)nu * sizeh + site, rec);
            }
        }
    }
}

enum PtcMode : int { kPtcApply = 0, kPtcFcg = 1, kPtcTrain = 2 };

// Layer 2: z(x) = w0 r(x) + sum_nu T_nu^dag g_nu (x), even x.
//  kPtcApply: out = z
//  kPtcFcg:   out = z, partials (Re<z,q>, Re<z,r>)        (aux = q; q == nullptr -> 0)
//  kPtcTrain: out = e = z - x, partials (|e|^2, Re<e,r>)  (aux = x)
template <int MODE, bool SGN, typename TR>
__global__ __launch_bounds__(kBlock) void k_ptc_layer2(PtcGeom A, const b6* __restrict__ g,
                                                       const TR* __restrict__ r, TR* __restrict__ out,
                                                       const TR* __restrict__ aux, float w0, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const int x = blockIdx.x * kBlock + threadIdx.x;
    const int sizeh = A.g.sizeh;
    double v0 = 0.0, v1 = 0.0;
    if (x < sizeh) {
        const Nbrs nb = site_neighbours<0>(A.g, x);
        float2 acc[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
#pragma unroll
        for (int mu = 0; mu < 4; mu++) {
            // adjoint of the forward transport: U_mu(x - mu)^dag e^{-i th} g_mu(x - mu)
            const b6 ga = ldg(g + (size_t)mu * sizeh + nb.im[mu]);
            const V3 a = xport<true, SGN>(A, 2 * mu + 1, nb.im[mu], bf_unpack(ga.v[0]), bf_unpack(ga.v[1]),
                                          bf_unpack(ga.v[2]));
            // adjoint of the backward transport: U_mu(x) e^{i th} g_{mu+4}(x + mu)
            const b6 gb = ldg(g + (size_t)(mu + 4) * sizeh + nb.ip[mu]);
            const V3 b = xport<false, SGN>(A, 2 * mu, x, bf_unpack(gb.v[0]), bf_unpack(gb.v[1]), bf_unpack(gb.v[2]));
#pragma unroll
            for (int c = 0; c < 3; c++) acc[c] = cadd(acc[c], cadd(a.c[c], b.c[c]));
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = (long)c * sizeh + x;
            const double2 rr = to_d2(r[k]);
            double2 z = make_double2((double)w0 * rr.x + (double)acc[c].x, (double)w0 * rr.y + (double)acc[c].y);
            if (MODE == kPtcTrain) {
                const double2 xx = to_d2(aux[k]);
                z = make_double2(z.x - xx.x, z.y - xx.y);
                v0 += z.x * z.x + z.y * z.y;
                v1 += z.x * rr.x + z.y * rr.y;
            } else if (MODE == kPtcFcg) {
                // the dots of the stored (rounded) z, as the next kernels see it
                z = to_d2(from_d2<TR>(z));
                if (aux) {
                    const double2 q = to_d2(aux[k]);
                    v0 += z.x * q.x + z.y * q.y;
                }
                v1 += z.x * rr.x + z.y * rr.y;
            }
            out[k] = from_d2<TR>(z);
        }
    }
    if (MODE != kPtcApply) block_partial<2>(v0, v1, ctl);
}

// Gradient Gram matrix: per odd site, P[nu][mu] += <T_nu e, T_mu b>; block
// partials of the 128 doubles (re/im of 64 entries), nu-major.
template <bool SGN>
__global__ __launch_bounds__(kBlock) void k_ptc_gram(PtcGeom A, const double2* __restrict__ b,
                                                     const double2* __restrict__ e, double* __restrict__ partials) {
    __shared__ double lds[kBlock / 64][128];
    const int y = blockIdx.x * kBlock + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    V3 t[8];
    const bool valid = y < A.g.sizeh;
    Nbrs nb;
    if (valid) {
        transports<SGN>(A, b, y, t);
        nb = site_neighbours<1>(A.g, y);
    }
#pragma unroll
    for (int nu = 0; nu < 8; nu++) {
        V3 a;
        if (valid) {
            a = (nu < 4) ? xport<false, SGN>(A, 2 * nu + 1, y, e, A.g.sizeh, nb.ip[nu])
                         : xport<true, SGN>(A, 2 * (nu - 4), nb.im[nu - 4], e, A.g.sizeh, nb.im[nu - 4]);
        }
#pragma unroll
        for (int mu = 0; mu < 8; mu++) {
            double pr = 0.0, pi = 0.0;
            if (valid) {
#pragma unroll
                for (int c = 0; c < 3; c++) {  // conj(a) * t
                    pr += (double)a.c[c].x * t[mu].c[c].x + (double)a.c[c].y * t[mu].c[c].y;
                    pi += (double)a.c[c].x * t[mu].c[c].y - (double)a.c[c].y * t[mu].c[c].x;
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                pr += __shfl_xor(pr, off, 64);
                pi += __shfl_xor(pi, off, 64);
            }
            if (lane == 0) {
                lds[wave][2 * (nu * 8 + mu)] = pr;
                lds[wave][2 * (nu * 8 + mu) + 1] = pi;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 128) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; w++) s += lds[w][threadIdx.x];
        partials[(size_t)blockIdx.x * 128 + threadIdx.x] = s;
    }
}

// sum the gram partials over blocks: one block per entry, each thread a
// strided subset of the blocks, then the block's fixed tree (deterministic)
__global__ __launch_bounds__(kBlock) void k_ptc_gram_sum(const double* __restrict__ partials, int nblocks,
                                                         double* __restrict__ out) {
    __shared__ double lds[kBlock / 64];
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) s += partials[(size_t)b * 128 + blockIdx.x];
    s = block_sum<kBlock>(s, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// ------------------------------------------------------------------ host side
static hipStream_t S() { return ctx().stream; }

static bool params_of(const float* p, PtcParams& W) {
    if (!p) return fail("ptc: null parameters");
    W.w0 = p[0];
    std::memcpy(W.w, p + 1, sizeof(W.w));
    return true;
}

// library-owned work space: g (8 odd BF16 channel fields), the FP16 link
// records, e, f, h (FP64), gram partials + sum, the link check flag
struct PtcScratch {
    void* p = nullptr;
    size_t bytes = 0;
    b6* g = nullptr;
    h8* la = nullptr;
    h6* lb = nullptr;
    double2* e = nullptr;
    double2* f = nullptr;
    double2* h = nullptr;
    double* part = nullptr;
    double* gram = nullptr;
    int* flag = nullptr;
};
static PtcScratch g_ps;

static bool ptc_scratch() {
    const long sh = ctx().lat.sizeh;
    const int nb = blocks_for(sh);
    const size_t gb = (size_t)8 * 12 * sh, lab = (size_t)8 * 16 * sh, lbb = (size_t)8 * 12 * sh,
                 eb = (size_t)48 * sh, pb = (size_t)nb * 128 * 8;
    const size_t need = gb + lab + lbb + 3 * eb + pb + 128 * 8 + 256 + 10 * 256;
    if (g_ps.bytes < need) {
        if (g_ps.p) (void)hipFree(g_ps.p);
        g_ps = PtcScratch{};
        if (!hip_ok(hipMalloc(&g_ps.p, need), "ptc scratch")) return false;
        g_ps.bytes = need;
    }
    char* c = (char*)g_ps.p;
    auto take = [&](size_t n) {
        void* q = c;
        c += (n + 255) & ~(size_t)255;
        return q;
    };
    g_ps.g = (b6*)take(gb);
    g_ps.la = (h8*)take(lab);
    g_ps.lb = (h6*)take(lbb);
    g_ps.e = (double2*)take(eb);
    g_ps.f = (double2*)take(eb);
    g_ps.h = (double2*)take(eb);
    g_ps.part = (double*)take(pb);
    g_ps.gram = (double*)take(128 * 8);
    g_ps.flag = (int*)take(256);
    return true;
}

// Links of the preconditioner: FP16 records built from the FP32 link copy and
// the FP32 phases (the compact 24 B form when every phase is +-1, reported in
// sgn); needs the scratch.  Returns false on error.
static bool geom_of(const void* u_f, const void* ph_f, PtcGeom& A, bool& sgn, const char* nm) {
    const Lattice& L = ctx().lat;
    const char* u = (const char*)resolve(u_f, 8 * L.su3f_bytes(), nm);
    const char* ph = (const char*)resolve(ph_f, 8 * L.fsoa_bytes(), nm);
    if (!u || !ph) return false;
    const long sh = L.sizeh;
    PtcLinkSrc src;
    for (int l = 0; l < 8; l++) {
        src.u[l] = (const float2*)(u + l * L.su3f_bytes());
        src.ph[l] = (const float*)(ph + l * L.fsoa_bytes());
        src.la[l] = g_ps.la + (size_t)l * sh;
        src.lb[l] = g_ps.lb + (size_t)l * sh;  // read as an h4 array when sgn (8 B records)
        A.la[l] = src.la[l];
        A.lb[l] = src.lb[l];
    }
    src.sizeh = (int)sh;
    const dim3 grid((unsigned)((8 * sh + kBlock - 1) / kBlock));
    int flag = 0;
    if (!hip_ok(hipMemsetAsync(g_ps.flag, 0, sizeof(int), S()), nm)) return false;
    hipLaunchKernelGGL(k_ptc_links_check, grid, dim3(kBlock), 0, S(), src, g_ps.flag);
    if (!hip_ok(hipGetLastError(), nm) || ostp_memcpy_dtoh(&flag, g_ps.flag, sizeof flag) != 0) return false;
    sgn = flag == 0;
    hipLaunchKernelGGL(k_ptc_links, grid, dim3(kBlock), 0, S(), src, sgn ? 1 : 0);
    A.g.nd0 = L.nd[0];
    A.g.nd0h = L.nd[0] / 2;
    A.g.nd1 = L.nd[1];
    A.g.nd2 = L.nd[2];
    A.g.nd3 = L.nd[3];
    A.g.sizeh = (int)sh;
    return hip_ok(hipGetLastError(), nm);
}

// M = W^dag W in the real 16 x 16 form of the layer-1 MFMA
static PtcMix mix_of(const PtcParams& W) {
    PtcMix M;
    for (int nu = 0; nu < 8; nu++)
        for (int mu = 0; mu < 8; mu++) {
            double a = 0.0, b = 0.0;  // sum_c conj(w[c][nu]) w[c][mu]
            for (int c = 0; c < kC; c++) {
                const double xr = W.w[c][nu][0], xi = W.w[c][nu][1], yr = W.w[c][mu][0], yi = W.w[c][mu][1];
                a += xr * yr + xi * yi;
                b += xr * yi - xi * yr;
            }
            M.m[2 * nu][2 * mu] = (float)a;
            M.m[2 * nu][2 * mu + 1] = (float)-b;
            M.m[2 * nu + 1][2 * mu] = (float)b;
            M.m[2 * nu + 1][2 * mu + 1] = (float)a;
        }
    return M;
}

template <bool SGN, typename TR>
static void launch_apply_t(const PtcGeom& A, const PtcMix& M, float w0, const TR* r, TR* z, int mode,
                           const TR* aux, const SolverCtl& ctl) {
    const long sh = A.g.sizeh;
    hipLaunchKernelGGL((k_ptc_layer1<SGN, TR>), dim3((unsigned)((sh + 255) / 256)), dim3(256), 0, S(), A, r, g_ps.g,
                       M, ctl);
    const int nb = blocks_for(sh);
    const b6* g = g_ps.g;
    if (mode == kPtcApply)
        hipLaunchKernelGGL((k_ptc_layer2<kPtcApply, SGN, TR>), dim3(nb), dim3(kBlock), 0, S(), A, g, r, z, aux, w0,
                           ctl);
    else if (mode == kPtcFcg)
        hipLaunchKernelGGL((k_ptc_layer2<kPtcFcg, SGN, TR>), dim3(nb), dim3(kBlock), 0, S(), A, g, r, z, aux, w0, ctl);
    else
        hipLaunchKernelGGL((k_ptc_layer2<kPtcTrain, SGN, TR>), dim3(nb), dim3(kBlock), 0, S(), A, g, r, z, aux, w0,
                           ctl);
}

// z = N r on device pointers (r, z even fields, FP64 or FP32), queued on the library stream
template <typename TR>
static bool launch_apply(const PtcGeom& A, bool sgn, const PtcMix& M, float w0, const TR* r, TR* z, int mode,
                         const TR* aux, const SolverCtl& ctl) {
    if (sgn)
        launch_apply_t<true, TR>(A, M, w0, r, z, mode, aux, ctl);
    else
        launch_apply_t<false, TR>(A, M, w0, r, z, mode, aux, ctl);
    return hip_ok(hipGetLastError(), "ptc launch");
}

// CG pieces shared with solvers.hip
bool launch_cg_xr(void* x, void* r, const void* p, const void* s, const SolverCtl& ctl);
bool launch_cg_p(void* p, const void* r, const SolverCtl& ctl);
bool launch_cg_xr_f(void* x, void* r, const void* p, const void* s, const SolverCtl& ctl);
bool launch_cg_p_f(void* p, const void* r, const SolverCtl& ctl);
bool launch_add_f_to_d(void* x, const void* e);
bool run_cg_batches(const std::function<bool(int)>& enqueue, int first_batch);
int cg_solve_from(const void* u, const void* ph, const void* b, void* x, void* r, void* h, void* s, void* p,
                  double mass, double res, int max_cg, int* cg_out, double* true_res);

}  // namespace ostp

using namespace ostp;

extern "C" {

int ostp_ptc_apply(const float* params, const su3_soa_f* u_f, const float_soa* phases_f, const vec3_soa* in,
                   vec3_soa* out) {
    const char* nm = "ostp_ptc_apply";
    if (!require_init(nm)) return 1;
    PtcParams W;
    PtcGeom A;
    bool sgn = false;
    if (!params_of(params, W) || !ptc_scratch() || !geom_of(u_f, phases_f, A, sgn, nm)) return 1;
    const size_t vb = ctx().lat.vec3_bytes();
    const double2* r = (const double2*)resolve(in, vb, nm);
    double2* z = (double2*)resolve(out, vb, nm);
    if (!r || !z) return 1;
    if (!launch_apply<double2>(A, sgn, mix_of(W), W.w0, r, z, kPtcApply, nullptr, plain_ctl())) return 1;
    return entry_done(nm) ? 0 : 1;
}

double ostp_ptc_loss_grad(const float* params, const su3_soa* u, const su3_soa_f* u_f, ferm_param* pars,
                          const vec3_soa* b, const vec3_soa* x, int a_norm, double* grad) {
    const char* nm = "ostp_ptc_loss_grad";
    if (!require_init(nm)) return -1.0;
    if (!pars) {
        fail("%s: null ferm_param", nm);
        return -1.0;
    }
    PtcParams W;
    PtcGeom A;
    bool sgn = false;
    if (!params_of(params, W) || !ptc_scratch() || !geom_of(u_f, pars->phases_f, A, sgn, nm)) return -1.0;
    const Lattice& Lt = ctx().lat;
    const size_t vb = Lt.vec3_bytes();
    const double2* db = (const double2*)resolve(b, vb, nm);
    const double2* dx = (const double2*)resolve(x, vb, nm);
    const void* du = a_norm ? resolve(u, 8 * Lt.su3_bytes(), nm) : nullptr;
    const void* dph = a_norm ? resolve(pars->phases, 8 * Lt.dsoa_bytes(), nm) : nullptr;
    if (!db || !dx || (a_norm && (!du || !dph))) return -1.0;
    const SolverCtl ctl = plain_ctl();
    const long sh = A.g.sizeh;
    const int nb = blocks_for(sh);
    SolverState& st = *ctx().state_host;
    // e = N b - x ; |e|^2 and Re<e, b>
    if (!launch_apply(A, sgn, mix_of(W), W.w0, db, g_ps.e, kPtcTrain, dx, ctl) || !launch_finalize(kFinStore2, ctl, nb) || !pull_state())
        return -1.0;
    double L = st.sum, s0 = st.sum_im;
    const double2* adj = g_ps.e;  // dL = 2 Re<adj, dz>
    if (a_norm) {
        // L = <e, A e>, adj = f = A e, s0 = Re<f, b>
        const double m2 = pars->ferm_mass * pars->ferm_mass;
        if (!launch_mdagm<double>(du, dph, g_ps.e, g_ps.h, g_ps.f, m2, 2, kFinStore, ctl) || !pull_state())
            return -1.0;
        L = st.sum;
        if (!reduce_to_host(1, g_ps.f, db, 0, &s0, nullptr)) return -1.0;
        adj = g_ps.f;
    }
    if (sgn)
        hipLaunchKernelGGL((k_ptc_gram<true>), dim3(nb), dim3(kBlock), 0, S(), A, db, adj, g_ps.part);
    else
        hipLaunchKernelGGL((k_ptc_gram<false>), dim3(nb), dim3(kBlock), 0, S(), A, db, adj, g_ps.part);
    hipLaunchKernelGGL(k_ptc_gram_sum, dim3(128), dim3(kBlock), 0, S(), (const double*)g_ps.part, nb, g_ps.gram);
    double P[128];
    if (!hip_ok(hipGetLastError(), nm) || ostp_memcpy_dtoh(P, g_ps.gram, sizeof P) != 0) return -1.0;
    if (grad) {
        grad[0] = 2.0 * s0;
        for (int c = 0; c < kC; c++)
            for (int nu = 0; nu < 8; nu++) {
                // S = sum_mu conj(w[c][mu]) (conj(P[nu][mu]) + P[mu][nu])
                double sr = 0.0, si = 0.0;
                for (int mu = 0; mu < 8; mu++) {
                    const double wr = W.w[c][mu][0], wi = -W.w[c][mu][1];       // conj(w)
                    const double qr = P[2 * (nu * 8 + mu)] + P[2 * (mu * 8 + nu)];
                    const double qi = -P[2 * (nu * 8 + mu) + 1] + P[2 * (mu * 8 + nu) + 1];
                    sr += wr * qr - wi * qi;
                    si += wr * qi + wi * qr;
                }
                grad[1 + 2 * (8 * c + nu)] = 2.0 * sr;   // 2 conj(S)
                grad[2 + 2 * (8 * c + nu)] = -2.0 * si;
            }
    }
    return L;
}

int ostp_fcg_solve(const su3_soa* u, ferm_param* pars, const su3_soa_f* u_f, const float* params,
                   vec3_soa* solution, const vec3_soa* in, double res, int max_cg, int* cg_return,
                   int* fallback_cg, double* true_res) {
    const char* nm = "ostp_fcg_solve";
    if (!require_init(nm)) return INVERTER_FAILURE;
    if (!pars || !cg_return) {
        fail("%s: null argument", nm);
        return INVERTER_FAILURE;
    }
    const Lattice& L = ctx().lat;
    PtcParams W;
    PtcGeom A;
    bool sgn = false;
    if (!params_of(params, W) || !ptc_scratch() || !geom_of(u_f, pars->phases_f, A, sgn, nm))
        return INVERTER_FAILURE;
    const PtcMix M = mix_of(W);
    const size_t vb = L.vec3_bytes();
    const void* du = resolve(u, 8 * L.su3_bytes(), nm);
    const void* dph = resolve(pars->phases, 8 * L.dsoa_bytes(), nm);
    const double2* b = (const double2*)resolve(in, vb, nm);
    double2* x = (double2*)resolve(solution, vb, nm);
    if (!du || !dph || !b || !x) return INVERTER_FAILURE;
    // FCG work vectors r, h, q, p, z (library scratch, FP64)
    static void* work = nullptr;
    static size_t work_bytes = 0;
    if (work_bytes < 5 * vb) {
        if (work) (void)hipFree(work);
        work = nullptr;
        work_bytes = 0;
        if (!hip_ok(hipMalloc(&work, 5 * vb), "fcg work")) return INVERTER_FAILURE;
        work_bytes = 5 * vb;
    }
    char* wv = (char*)work;
    double2 *r = (double2*)wv, *h = (double2*)(wv + vb), *q = (double2*)(wv + 2 * vb), *p = (double2*)(wv + 3 * vb),
            *z = (double2*)(wv + 4 * vb);
    const double m2 = pars->ferm_mass * pars->ferm_mass;
    const long n3 = 3 * L.sizeh;
    const int nbh = blocks_for(L.sizeh);
    double sn = 0.0;
    if (!reduce_to_host(0, b, nullptr, 0, &sn, nullptr)) return INVERTER_FAILURE;
    SolverState& hs = *ctx().state_host;
    std::memset(&hs, 0, sizeof hs);
    hs.res = res;
    hs.safety = 0.95;
    hs.max_cg = max_cg;
    hs.source_norm = sn;
    hs.status = kRunning;
    if (!push_state()) return INVERTER_FAILURE;
    auto ctl_pred = [](int pred) {
        SolverCtl c = plain_ctl();
        c.pred = pred;
        return c;
    };
    // r = b - A x ; z = N r ; <z, r> ; p = z
    if (!launch_mdagm<double>(du, dph, x, h, q, m2, 1, kFinStore, ctl_pred(kAlways)) || !launch_sub(b, q, r))
        return INVERTER_FAILURE;
    if (!launch_apply<double2>(A, sgn, M, W.w0, r, z, kPtcFcg, nullptr, ctl_pred(kAlways)) ||
        !launch_finalize(kFinFcgInit, ctl_pred(kAlways), nbh) ||
        !hip_ok(hipMemcpyAsync(p, z, vb, hipMemcpyDeviceToDevice, S()), "p = z"))
        return INVERTER_FAILURE;
    auto iteration = [&](int) {
        const SolverCtl run = ctl_pred(kIfRunning);
        // q = A p, alpha = <z,r> / <p,q> ; x += alpha p ; r -= alpha q ; |r|^2, exit test ;
        // z = N r ; beta = -alpha <z,q> / <z,r>_old ; p = z + beta p
        return launch_mdagm<double>(du, dph, p, h, q, m2, 2, kFinFcgAlpha, run) && launch_cg_xr(x, r, p, q, run) &&
               launch_finalize(kFinFcgRes, run, nbh) && launch_apply(A, sgn, M, W.w0, r, z, kPtcFcg, q, run) &&
               launch_finalize(kFinFcgBeta, run, nbh) && launch_cg_p(p, z, run);
    };
    if (!run_cg_batches(iteration, 4)) return INVERTER_FAILURE;
    int fcg_its = hs.cg;
    // FP64 true residual
    if (!launch_mdagm<double>(du, dph, x, h, q, m2, 1, kFinStore, ctl_pred(kAlways)) || !launch_sub(b, q, r))
        return INVERTER_FAILURE;
    double rr = 0.0;
    if (!reduce_to_host(0, r, nullptr, 0, &rr, nullptr)) return INVERTER_FAILURE;
    double tr = std::sqrt(rr / sn);
    int extra = 0;
    int status = tr <= res ? INVERTER_SUCCESS : INVERTER_FAILURE;
    if (status != INVERTER_SUCCESS) {
        // plain CG from the current iterate (ker_invert_openacc semantics)
        if (verbosity_lv > 0)
            printf("ostp_fcg_solve: flexible CG stopped at %e after %d iterations, plain CG takes over\n", tr, fcg_its);
        status = cg_solve_from(du, dph, b, x, r, h, q, p, pars->ferm_mass, res, max_cg, &extra, &tr);
    }
    *cg_return = fcg_its;
    if (fallback_cg) *fallback_cg = extra;
    if (true_res) *true_res = tr;
    (void)n3;
    return status;
}

// Flexible CG with the network in FP32 (links, phases, vectors and the
// network's input/output), with FP64 reliable updates: the FP32 iteration
// accumulates its update e; whenever its residual has dropped by inner_res
// since the last update, x += e in FP64, the FP64 true residual r = b - A x
// replaces the FP32 one, and the iteration continues along the same search
// direction (z = N r, beta from the kept q and <z,r>).  The exit test is the
// FP64 true residual |b - A x| / |b| <= res.  An update that leaves more than
// (1 + inner_res) / 2 of the previous FP64 residual (or an exhausted
// iteration budget) hands the iterate to the FP64 plain CG, as
// ostp_fcg_solve does.
int ostp_fcg_mixed_solve(const su3_soa* u, ferm_param* pars, const su3_soa_f* u_f, const float* params,
                         vec3_soa* solution, const vec3_soa* in, double res, int max_cg, double inner_res,
                         int* cg_return, int* outer_return, int* fallback_cg, double* true_res) {
    const char* nm = "ostp_fcg_mixed_solve";
    if (!require_init(nm)) return INVERTER_FAILURE;
    if (!pars || !cg_return) {
        fail("%s: null argument", nm);
        return INVERTER_FAILURE;
    }
    if (!(inner_res > 0.0 && inner_res < 1.0)) {
        fail("%s: inner tolerance %g outside (0, 1)", nm, inner_res);
        return INVERTER_FAILURE;
    }
    const Lattice& L = ctx().lat;
    PtcParams W;
    PtcGeom A;
    bool sgn = false;
    if (!params_of(params, W) || !ptc_scratch() || !geom_of(u_f, pars->phases_f, A, sgn, nm))
        return INVERTER_FAILURE;
    const PtcMix M = mix_of(W);
    const size_t vb = L.vec3_bytes(), vbf = L.vec3f_bytes();
    const void* du = resolve(u, 8 * L.su3_bytes(), nm);
    const void* dph = resolve(pars->phases, 8 * L.dsoa_bytes(), nm);
    const void* duf = resolve(u_f, 8 * L.su3f_bytes(), nm);
    const void* dphf = resolve(pars->phases_f, 8 * L.fsoa_bytes(), nm);
    const double2* b = (const double2*)resolve(in, vb, nm);
    double2* x = (double2*)resolve(solution, vb, nm);
    if (!du || !dph || !duf || !dphf || !b || !x) return INVERTER_FAILURE;
    // work vectors: FP64 r, h, q, p (p only for the fallback CG); FP32 r, h, q, p, z, e
    static void* work = nullptr;
    static size_t work_bytes = 0;
    const size_t need = 4 * vb + 6 * vbf;
    if (work_bytes < need) {
        if (work) (void)hipFree(work);
        work = nullptr;
        work_bytes = 0;
        if (!hip_ok(hipMalloc(&work, need), "fcg work")) return INVERTER_FAILURE;
        work_bytes = need;
    }
    char* wv = (char*)work;
    double2 *r = (double2*)wv, *h = (double2*)(wv + vb), *q = (double2*)(wv + 2 * vb), *p = (double2*)(wv + 3 * vb);
    char* wf = wv + 4 * vb;
    float2 *rf = (float2*)wf, *hf = (float2*)(wf + vbf), *qf = (float2*)(wf + 2 * vbf), *pf = (float2*)(wf + 3 * vbf),
           *zf = (float2*)(wf + 4 * vbf), *ef = (float2*)(wf + 5 * vbf);
    const double mass = pars->ferm_mass, m2 = mass * mass;
    const int nbh = blocks_for(L.sizeh);
    const long n3 = 3 * L.sizeh;
    double sn = 0.0;
    if (!reduce_to_host(0, b, nullptr, 0, &sn, nullptr)) return INVERTER_FAILURE;
    SolverState& hs = *ctx().state_host;
    auto ctl_pred = [](int pred) {
        SolverCtl c = plain_ctl();
        c.pred = pred;
        return c;
    };
    // FP64 true residual r = b - A x, |r|^2
    auto true_residual = [&](double* rr) {
        return launch_mdagm<double>(du, dph, x, h, q, m2, 1, kFinStore, ctl_pred(kAlways)) && launch_sub(b, q, r) &&
               reduce_to_host(0, r, nullptr, 0, rr, nullptr);
    };
    auto iteration = [&](int) {
        const SolverCtl run = ctl_pred(kIfRunning);
        return launch_mdagm<float>(duf, dphf, pf, hf, qf, m2, 2, kFinFcgAlpha, run) &&
               launch_cg_xr_f(ef, rf, pf, qf, run) && launch_finalize(kFinFcgRes, run, nbh) &&
               launch_apply<float2>(A, sgn, M, W.w0, rf, zf, kPtcFcg, qf, run) &&
               launch_finalize(kFinFcgBeta, run, nbh) && launch_cg_p_f(pf, zf, run);
    };
    double rr = 0.0;
    if (!true_residual(&rr)) return INVERTER_FAILURE;
    int outer = 0;
    bool stalled = false;
    std::memset(&hs, 0, sizeof hs);
    hs.res = inner_res;
    hs.safety = 1.0;
    hs.max_cg = max_cg;
    while (std::sqrt(rr / sn) > res && hs.cg < max_cg && !stalled) {
        // the FP32 iteration restarts from the FP64 residual: e = 0, r_f = r;
        // the first time z = N r, p = z; afterwards z = N r, beta, p = z + beta p
        hs.source_norm = rr;
        hs.status = kRunning;
        if (!push_state() || !launch_d2f(r, rf, n3) || !hip_ok(hipMemsetAsync(ef, 0, vbf, S()), "e = 0"))
            return INVERTER_FAILURE;
        if (outer == 0) {
            if (!launch_apply<float2>(A, sgn, M, W.w0, rf, zf, kPtcFcg, nullptr, ctl_pred(kAlways)) ||
                !launch_finalize(kFinFcgInit, ctl_pred(kAlways), nbh) ||
                !hip_ok(hipMemcpyAsync(pf, zf, vbf, hipMemcpyDeviceToDevice, S()), "p = z"))
                return INVERTER_FAILURE;
        } else if (!launch_apply<float2>(A, sgn, M, W.w0, rf, zf, kPtcFcg, qf, ctl_pred(kAlways)) ||
                   !launch_finalize(kFinFcgBeta, ctl_pred(kAlways), nbh) || !launch_cg_p_f(pf, zf, ctl_pred(kAlways)))
            return INVERTER_FAILURE;
        if (!run_cg_batches(iteration, 4)) return INVERTER_FAILURE;  // leaves hs pulled
        outer++;
        const double rr_old = rr;
        if (!launch_add_f_to_d(x, ef) || !true_residual(&rr)) return INVERTER_FAILURE;
        const double keep = 0.5 * (1.0 + inner_res);  // what a working update leaves at most
        stalled = !(rr < keep * keep * rr_old);
        if (verbosity_lv > 1)
            printf("%s: update %d after %d FP32 iterations, FP64 residual %e\n", nm, outer, hs.cg,
                   std::sqrt(rr / sn));
    }
    const int total = hs.cg;
    double tr = std::sqrt(rr / sn);
    int extra = 0;
    int status = tr <= res ? INVERTER_SUCCESS : INVERTER_FAILURE;
    if (status != INVERTER_SUCCESS) {
        if (verbosity_lv > 0)
            printf("%s: stopped at %e after %d iterations and %d updates, plain CG takes over\n", nm, tr, total,
                   outer);
        status = cg_solve_from(du, dph, b, x, r, h, q, p, mass, res, max_cg, &extra, &tr);
    }
    *cg_return = total;
    if (outer_return) *outer_return = outer;
    if (fallback_cg) *fallback_cg = extra;
    if (true_res) *true_res = tr;
    return status;
}

}  // extern "C"
