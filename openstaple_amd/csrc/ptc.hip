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
// spectrum), as conjugate-gradient methods need.  Links and transports are FP32; the
// channel mixing (h = W t, g = W^dag h) runs on the matrix cores as two
// chained BF16 MFMAs with FP32 accumulation (32x32x8: the 32x32 result of the
// first product has exactly the lane layout of the B operand of the second,
// so h never leaves registers); the transported vectors reach the MFMA
// operand layout through LDS.
//
// Training kernels: the loss |e|^2 or <e, A e> (e = N b - x) over (b, x)
// pairs, its gradient via the 8x8 Gram matrix P[nu][mu] = <T_nu a, T_mu b>
// with a = e or A e.  The CPU
// oracle (oracle/ptc.c) restates forward, loss and gradient in FP64.
#include "ostp_dslash.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace ostp {

constexpr int kC = OSTP_PTC_CHANNELS;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ __forceinline__ short to_bf16(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

struct PtcParams {
    float w0;
    float w[kC][8][2];  // (re, im)
};

struct PtcGeom {
    const float2* u[8];
    const float* ph[8];
    SiteGeom g;
};

__device__ __forceinline__ float2 ldv(const double2* p) {
    const double2 v = *p;
    return make_float2((float)v.x, (float)v.y);
}
__device__ __forceinline__ float2 ldv(const float2* p) { return *p; }

struct V3 {
    float2 c[3];
};

// One transport (FP32): U(iu) e^{i th(iu)} in(iv)  (DAG = false)
//                       U(iu)^dag e^{-i th(iu)} in(iv)  (DAG = true)
template <bool DAG, typename In>
__device__ __forceinline__ V3 xport(const float2* __restrict__ U, const float* __restrict__ P,
                                    const In* __restrict__ in, int sizeh, int iu, int iv) {
    const float2 a0 = U[iu], a1 = U[sizeh + iu], a2 = U[2 * sizeh + iu];
    const float2 b0 = U[3 * sizeh + iu], b1 = U[4 * sizeh + iu], b2 = U[5 * sizeh + iu];
    float s, c;
    sincos_t(P[iu], &s, &c);
    const float2 ph = DAG ? make_float2(c, -s) : make_float2(c, s);
    const float2 v0 = cmul(ldv(in + iv), ph), v1 = cmul(ldv(in + sizeh + iv), ph), v2 = cmul(ldv(in + 2 * sizeh + iv), ph);
    const float2 x0 = csub(cmul(a1, b2), cmul(a2, b1));
    const float2 x1 = csub(cmul(a2, b0), cmul(a0, b2));
    const float2 x2 = csub(cmul(a0, b1), cmul(a1, b0));
    V3 o;
    if (!DAG) {
        float2 r0 = cmul(a0, v0); r0 = cfma(a1, v1, r0); o.c[0] = cfma(a2, v2, r0);
        float2 r1 = cmul(b0, v0); r1 = cfma(b1, v1, r1); o.c[1] = cfma(b2, v2, r1);
        float2 r2 = cmul(cconj(x0), v0); r2 = cfmac(x1, v1, r2); o.c[2] = cfmac(x2, v2, r2);
    } else {
        float2 r0 = cmul(cconj(a0), v0); r0 = cfmac(b0, v1, r0); o.c[0] = cfma(x0, v2, r0);
        float2 r1 = cmul(cconj(a1), v0); r1 = cfmac(b1, v1, r1); o.c[1] = cfma(x1, v2, r1);
        float2 r2 = cmul(cconj(a2), v0); r2 = cfmac(b2, v1, r2); o.c[2] = cfma(x2, v2, r2);
    }
    return o;
}

// The eight transports T_nu in (in: even field) at odd half-site y.
template <typename In>
__device__ __forceinline__ void transports(const PtcGeom& A, const In* in, int y, V3 t[8]) {
    const Nbrs nb = site_neighbours<1>(A.g, y);
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
        t[mu] = xport<false>(A.u[2 * mu + 1], A.ph[2 * mu + 1], in, A.g.sizeh, y, nb.ip[mu]);
        t[mu + 4] = xport<true>(A.u[2 * mu], A.ph[2 * mu], in, A.g.sizeh, nb.im[mu], nb.im[mu]);
    }
}

// Layer 1 + channel mixing: g_nu(y) = sum_c conj(w[c][nu]) sum_mu w[c][mu] T_mu r (y), odd y.
// One wave = 64 odd sites; MFMA columns n = colour * 64 + site-in-wave.
template <typename In>
__global__ __launch_bounds__(256) void k_ptc_layer1(PtcGeom A, const In* __restrict__ r, float2* __restrict__ gout,
                                                    PtcParams W, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    __shared__ bf16x4 lds[4][192 * 4];  // per wave: [n][k] with k = 2 nu + re/im (16 bf16 per column)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int base = blockIdx.x * 256 + wave * 64;
    const int y = base + lane;
    const int sizeh = A.g.sizeh;
    V3 t[8];
    if (y < sizeh) {
        transports(A, r, y, t);
    } else {
#pragma unroll
        for (int mu = 0; mu < 8; mu++) t[mu].c[0] = t[mu].c[1] = t[mu].c[2] = make_float2(0.f, 0.f);
    }
    bf16x4* T = lds[wave];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int n = c * 64 + lane;
#pragma unroll
        for (int q = 0; q < 4; q++) {  // k = 4q .. 4q+3 = (nu = 2q, re, im), (nu = 2q+1, re, im)
            bf16x4 v;
            v[0] = to_bf16(t[2 * q].c[c].x);
            v[1] = to_bf16(t[2 * q].c[c].y);
            v[2] = to_bf16(t[2 * q + 1].c[c].x);
            v[3] = to_bf16(t[2 * q + 1].c[c].y);
            T[n * 4 + q] = v;
        }
    }
    __syncthreads();
    // A operands.  Lane: row m = lane % 32, k = 8 s + 4 (lane / 32) + e.
    const int m = lane & 31, kh = 4 * (lane >> 5);
    bf16x4 a1[2], a2[3];
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            // stage 1: H = W t, rows m = 2c + ri (c < 12), cols k = 2 nu + rk
            const int k = 8 * s + kh + e, c = m >> 1, nu = k >> 1;
            float v = 0.f;
            if (c < kC) {
                const float wr = W.w[c][nu][0], wi = W.w[c][nu][1];
                v = (m & 1) ? ((k & 1) ? wr : wi) : ((k & 1) ? -wi : wr);
            }
            a1[s][e] = to_bf16(v);
        }
#pragma unroll
    for (int s = 0; s < 3; s++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            // stage 2: G = W^dag H, rows m = 2 nu + ri (nu < 8), cols k = 2c + rk (c < 12)
            const int k = 8 * s + kh + e, nu = m >> 1, c = k >> 1;
            float v = 0.f;
            if (nu < 8) {
                const float wr = W.w[c][nu][0], wi = W.w[c][nu][1];
                v = (m & 1) ? ((k & 1) ? wr : -wi) : ((k & 1) ? wi : wr);
            }
            a2[s][e] = to_bf16(v);
        }
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const int n = j * 32 + m;
        f32x16 acc1 = {0.f};
#pragma unroll
        for (int s = 0; s < 2; s++) acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1[s], T[n * 4 + 2 * s + (lane >> 5)], acc1, 0, 0, 0);
        f32x16 acc2 = {0.f};
#pragma unroll
        for (int s = 0; s < 3; s++) {
            bf16x4 b;
#pragma unroll
            for (int e = 0; e < 4; e++) b[e] = to_bf16(acc1[4 * s + e]);
            acc2 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a2[s], b, acc2, 0, 0, 0);
        }
        // lane holds column n, rows 8 (v/4) + 4 (lane/32) + v%4; rows < 16 are (nu, re/im)
        const int colour = n >> 6, site = base + (n & 63);
        if (site < sizeh) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = 8 * (q >> 1) + kh + 2 * (q & 1);
                const int nu = row >> 1;
                gout[(size_t)(nu * 3 + colour) * sizeh + site] = make_float2(acc2[2 * q], acc2[2 * q + 1]);
            }
        }
    }
}

enum PtcMode : int { kPtcApply = 0, kPtcFcg = 1, kPtcTrain = 2 };

// Layer 2: z(x) = w0 r(x) + sum_nu T_nu^dag g_nu (x), even x.
//  kPtcApply: out = z
//  kPtcFcg:   out = z, partials (Re<z,q>, Re<z,r>)        (aux = q; q == nullptr -> 0)
//  kPtcTrain: out = e = z - x, partials (|e|^2, Re<e,r>)  (aux = x)
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_ptc_layer2(PtcGeom A, const float2* __restrict__ g,
                                                       const double2* __restrict__ r, double2* __restrict__ out,
                                                       const double2* __restrict__ aux, float w0, SolverCtl ctl) {
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
            const V3 a = xport<true>(A.u[2 * mu + 1], A.ph[2 * mu + 1], g + (size_t)mu * 3 * sizeh, sizeh, nb.im[mu],
                                     nb.im[mu]);
            // adjoint of the backward transport: U_mu(x) e^{i th} g_{mu+4}(x + mu)
            const V3 b = xport<false>(A.u[2 * mu], A.ph[2 * mu], g + (size_t)(mu + 4) * 3 * sizeh, sizeh, x, nb.ip[mu]);
#pragma unroll
            for (int c = 0; c < 3; c++) acc[c] = cadd(acc[c], cadd(a.c[c], b.c[c]));
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long k = (long)c * sizeh + x;
            const double2 rr = r[k];
            double2 z = make_double2((double)w0 * rr.x + (double)acc[c].x, (double)w0 * rr.y + (double)acc[c].y);
            if (MODE == kPtcTrain) {
                const double2 xx = aux[k];
                z = make_double2(z.x - xx.x, z.y - xx.y);
                v0 += z.x * z.x + z.y * z.y;
                v1 += z.x * rr.x + z.y * rr.y;
            } else if (MODE == kPtcFcg) {
                if (aux) {
                    const double2 q = aux[k];
                    v0 += z.x * q.x + z.y * q.y;
                }
                v1 += z.x * rr.x + z.y * rr.y;
            }
            out[k] = z;
        }
    }
    if (MODE != kPtcApply) block_partial<2>(v0, v1, ctl);
}

// Gradient Gram matrix: per odd site, P[nu][mu] += <T_nu e, T_mu b>; block
// partials of the 128 doubles (re/im of 64 entries), nu-major.
__global__ __launch_bounds__(kBlock) void k_ptc_gram(PtcGeom A, const double2* __restrict__ b,
                                                     const double2* __restrict__ e, double* __restrict__ partials) {
    __shared__ double lds[kBlock / 64][128];
    const int y = blockIdx.x * kBlock + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    V3 t[8];
    const bool valid = y < A.g.sizeh;
    Nbrs nb;
    if (valid) {
        transports(A, b, y, t);
        nb = site_neighbours<1>(A.g, y);
    }
#pragma unroll
    for (int nu = 0; nu < 8; nu++) {
        V3 a;
        if (valid) {
            a = (nu < 4) ? xport<false>(A.u[2 * nu + 1], A.ph[2 * nu + 1], e, A.g.sizeh, y, nb.ip[nu])
                         : xport<true>(A.u[2 * (nu - 4)], A.ph[2 * (nu - 4)], e, A.g.sizeh, nb.im[nu - 4],
                                       nb.im[nu - 4]);
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

// sum the gram partials over blocks (one thread per entry, fixed order)
__global__ __launch_bounds__(128) void k_ptc_gram_sum(const double* __restrict__ partials, int nblocks,
                                                      double* __restrict__ out) {
    double s = 0.0;
    for (int b = 0; b < nblocks; b++) s += partials[(size_t)b * 128 + threadIdx.x];
    out[threadIdx.x] = s;
}

// ------------------------------------------------------------------ host side
static hipStream_t S() { return ctx().stream; }

static bool params_of(const float* p, PtcParams& W) {
    if (!p) return fail("ptc: null parameters");
    W.w0 = p[0];
    std::memcpy(W.w, p + 1, sizeof(W.w));
    return true;
}

static bool geom_of(const void* u_f, const void* ph_f, PtcGeom& A, const char* nm) {
    const Lattice& L = ctx().lat;
    const char* u = (const char*)resolve(u_f, 8 * L.su3f_bytes(), nm);
    const char* ph = (const char*)resolve(ph_f, 8 * L.fsoa_bytes(), nm);
    if (!u || !ph) return false;
    for (int l = 0; l < 8; l++) {
        A.u[l] = (const float2*)(u + l * L.su3f_bytes());
        A.ph[l] = (const float*)(ph + l * L.fsoa_bytes());
    }
    A.g.nd0 = L.nd[0];
    A.g.nd0h = L.nd[0] / 2;
    A.g.nd1 = L.nd[1];
    A.g.nd2 = L.nd[2];
    A.g.nd3 = L.nd[3];
    A.g.sizeh = (int)L.sizeh;
    return true;
}

// library-owned work space: g (8 odd FP32 fields), e, f, h (FP64), gram partials + sum
struct PtcScratch {
    void* p = nullptr;
    size_t bytes = 0;
    float2* g = nullptr;
    double2* e = nullptr;
    double2* f = nullptr;
    double2* h = nullptr;
    double* part = nullptr;
    double* gram = nullptr;
};
static PtcScratch g_ps;

static bool ptc_scratch() {
    const long sh = ctx().lat.sizeh;
    const int nb = blocks_for(sh);
    const size_t gb = (size_t)8 * 24 * sh, eb = (size_t)48 * sh, pb = (size_t)nb * 128 * 8;
    const size_t need = gb + 3 * eb + pb + 128 * 8 + 6 * 256;
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
    g_ps.g = (float2*)take(gb);
    g_ps.e = (double2*)take(eb);
    g_ps.f = (double2*)take(eb);
    g_ps.h = (double2*)take(eb);
    g_ps.part = (double*)take(pb);
    g_ps.gram = (double*)take(128 * 8);
    return true;
}

// z = N r on device pointers (r, z FP64 even fields), queued on the library stream
static bool launch_apply(const PtcGeom& A, const PtcParams& W, const double2* r, double2* z, int mode,
                         const double2* aux, const SolverCtl& ctl) {
    const long sh = A.g.sizeh;
    hipLaunchKernelGGL((k_ptc_layer1<double2>), dim3((sh + 255) / 256), dim3(256), 0, S(), A, r, g_ps.g, W, ctl);
    const int nb = blocks_for(sh);
    if (mode == kPtcApply)
        hipLaunchKernelGGL((k_ptc_layer2<kPtcApply>), dim3(nb), dim3(kBlock), 0, S(), A, (const float2*)g_ps.g, r, z,
                           aux, W.w0, ctl);
    else if (mode == kPtcFcg)
        hipLaunchKernelGGL((k_ptc_layer2<kPtcFcg>), dim3(nb), dim3(kBlock), 0, S(), A, (const float2*)g_ps.g, r, z,
                           aux, W.w0, ctl);
    else
        hipLaunchKernelGGL((k_ptc_layer2<kPtcTrain>), dim3(nb), dim3(kBlock), 0, S(), A, (const float2*)g_ps.g, r, z,
                           aux, W.w0, ctl);
    return hip_ok(hipGetLastError(), "ptc launch");
}

// CG pieces shared with solvers.hip
bool launch_cg_xr(void* x, void* r, const void* p, const void* s, const SolverCtl& ctl);
bool launch_cg_p(void* p, const void* r, const SolverCtl& ctl);
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
    if (!params_of(params, W) || !geom_of(u_f, phases_f, A, nm) || !ptc_scratch()) return 1;
    const size_t vb = ctx().lat.vec3_bytes();
    const double2* r = (const double2*)resolve(in, vb, nm);
    double2* z = (double2*)resolve(out, vb, nm);
    if (!r || !z) return 1;
    if (!launch_apply(A, W, r, z, kPtcApply, nullptr, plain_ctl())) return 1;
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
    if (!params_of(params, W) || !geom_of(u_f, pars->phases_f, A, nm) || !ptc_scratch()) return -1.0;
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
    if (!launch_apply(A, W, db, g_ps.e, kPtcTrain, dx, ctl) || !launch_finalize(kFinStore2, ctl, nb) || !pull_state())
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
    hipLaunchKernelGGL(k_ptc_gram, dim3(nb), dim3(kBlock), 0, S(), A, db, adj, g_ps.part);
    hipLaunchKernelGGL(k_ptc_gram_sum, dim3(1), dim3(128), 0, S(), (const double*)g_ps.part, nb, g_ps.gram);
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
    if (!params_of(params, W) || !geom_of(u_f, pars->phases_f, A, nm) || !ptc_scratch()) return INVERTER_FAILURE;
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
    if (!launch_apply(A, W, r, z, kPtcFcg, nullptr, ctl_pred(kAlways)) ||
        !launch_finalize(kFinFcgInit, ctl_pred(kAlways), nbh) ||
        !hip_ok(hipMemcpyAsync(p, z, vb, hipMemcpyDeviceToDevice, S()), "p = z"))
        return INVERTER_FAILURE;
    int last_cg = -1, batch = 4;
    for (;;) {
        if (!pull_state()) return INVERTER_FAILURE;
        if (hs.status != kRunning) break;
        if (hs.cg == last_cg) {
            fail("%s: solver state did not advance", nm);
            return INVERTER_FAILURE;
        }
        last_cg = hs.cg;
        for (int j = 0; j < batch; j++) {
            const SolverCtl run = ctl_pred(kIfRunning);
            // q = A p, alpha = <z,r> / <p,q>
            if (!launch_mdagm<double>(du, dph, p, h, q, m2, 2, kFinFcgAlpha, run)) return INVERTER_FAILURE;
            // x += alpha p ; r -= alpha q ; |r|^2, exit test
            if (!launch_cg_xr(x, r, p, q, run) || !launch_finalize(kFinFcgRes, run, nbh)) return INVERTER_FAILURE;
            // z = N r ; beta = -alpha <z,q> / <z,r>_old ; p = z + beta p
            if (!launch_apply(A, W, r, z, kPtcFcg, q, run) || !launch_finalize(kFinFcgBeta, run, nbh) ||
                !launch_cg_p(p, z, run))
                return INVERTER_FAILURE;
        }
        batch = batch < 64 ? 2 * batch : 64;
    }
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

}  // extern "C"
