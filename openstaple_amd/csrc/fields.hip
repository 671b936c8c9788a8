// Geometry tables, U(1) phase tables, field generators and a copy kernel.
//
//  ostp_compute_nnp_and_nnm  OpenAcc/geometry.c:14-100 (single rank)
//  calc_u1_phases            OpenAcc/backfield.c:20-187, KSphase* backfield.h:27-41
//  ostp_generate_*           rules of OpenAcc/random_assignement.c:29-250 on a
//                            counter-based Philox4x32-10 stream (Salmon et al., SC11)
#include "ostp_blas.h"
#include "ostp_su3.h"

#include <cmath>

namespace ostp {

struct GeomArgs {
    int nd[4];
    int map[4];
    int sizeh;
};

// --------------------------------------------------------------- nn tables
__global__ __launch_bounds__(kBlock) void k_nn_tables(int* __restrict__ nnp, int* __restrict__ nnm, GeomArgs g) {
    const int nd0 = g.nd[0], nd1 = g.nd[1], nd2 = g.nd[2], nd3 = g.nd[3];
    const long vol = (long)nd0 * nd1 * nd2 * nd3;
    const long site = (long)blockIdx.x * kBlock + threadIdx.x;
    if (site >= vol) return;
    int d0 = (int)(site % nd0);
    long t = site / nd0;
    int d1 = (int)(t % nd1);
    t /= nd1;
    int d2 = (int)(t % nd2);
    int d3 = (int)(t / nd2);
    const int vol1 = nd0, vol2 = nd1 * vol1, vol3 = nd2 * vol2;
    auto snum = [&](int a, int b, int c, int d) { return (a + b * vol1 + c * vol2 + d * vol3) / 2; };
    const int idxh = snum(d0, d1, d2, d3);
    const int parity = (d0 + d1 + d2 + d3) % 2;
    const int d0m = d0 == 0 ? nd0 - 1 : d0 - 1, d1m = d1 == 0 ? nd1 - 1 : d1 - 1;
    const int d2m = d2 == 0 ? nd2 - 1 : d2 - 1, d3m = d3 == 0 ? nd3 - 1 : d3 - 1;
    const int d0p = d0 + 1 == nd0 ? 0 : d0 + 1, d1p = d1 + 1 == nd1 ? 0 : d1 + 1;
    const int d2p = d2 + 1 == nd2 ? 0 : d2 + 1, d3p = d3 + 1 == nd3 ? 0 : d3 + 1;
    int* m = nnm + (long)idxh * 8;
    int* p = nnp + (long)idxh * 8;
    m[0 + parity] = snum(d0m, d1, d2, d3);
    m[2 + parity] = snum(d0, d1m, d2, d3);
    m[4 + parity] = snum(d0, d1, d2m, d3);
    m[6 + parity] = snum(d0, d1, d2, d3m);
    p[0 + parity] = snum(d0p, d1, d2, d3);
    p[2 + parity] = snum(d0, d1p, d2, d3);
    p[4 + parity] = snum(d0, d1, d2p, d3);
    p[6 + parity] = snum(d0, d1, d2, d3p);
}

// ------------------------------------------------------------------ phases
// The double operations below are the reference's, in its order and without
// contraction, so the table is bit-identical to the host computation.
struct BfArgs {
    double ex, ey, ez, bx, by, bz;
    double im_chem_pot, ferm_charge;
};

__device__ __forceinline__ double rebound_2pi(double a) {
#pragma clang fp contract(off)
    while (a > 0.5) a -= 1.0;
    while (a < -0.5) a += 1.0;
    return a * (2 * 3.14159265358979323846);
}

// bounded: calc_u1_phases (reduced to [-1/2, 1/2] turns, times 2 pi);
// otherwise calc_u1_phases_unb_no2pi (backfield.c:20-160): turns, unreduced
__global__ __launch_bounds__(kBlock) void k_u1_phases(double* __restrict__ phases, GeomArgs g, BfArgs bf,
                                                      int bounded) {
#pragma clang fp contract(off)
    const int nd0 = g.nd[0], nd1 = g.nd[1], nd2 = g.nd[2], nd3 = g.nd[3];
    const long vol = (long)nd0 * nd1 * nd2 * nd3;
    const long site = (long)blockIdx.x * kBlock + threadIdx.x;
    if (site >= vol) return;
    int d[4];
    d[0] = (int)(site % nd0);
    long t = site / nd0;
    d[1] = (int)(t % nd1);
    t /= nd1;
    d[2] = (int)(t % nd2);
    d[3] = (int)(t / nd2);
    const int idxh = (d[0] + d[1] * nd0 + d[2] * nd0 * nd1 + d[3] * nd0 * nd1 * nd2) / 2;
    const int xmap = g.map[0], ymap = g.map[1], zmap = g.map[2], tmap = g.map[3];
    const int x = d[xmap], tnx = g.nd[xmap];
    const int y = d[ymap], tny = g.nd[ymap];
    const int z = d[zmap], tnz = g.nd[zmap];
    const int tt = d[tmap], tnt = g.nd[tmap];
    const int parity = (x + y + z + tt) % 2;
    const int X = x + 1, Y = y + 1, Z = z + 1, T = tt + 1;
    const double chpotphase = bf.im_chem_pot / tnt;
    const long sh = g.sizeh;
    double arg;

    arg = (z - tnz / 2 + 1) * bf.by / (tnz * tnx);
    if (X == tnx) {
        arg -= (y - tny / 2 + 1) * tnx * bf.bz / (tnx * tny);
        arg -= (tt - tnt / 2 + 1) * tnx * bf.ex / (tnx * tnt);
    }
    arg *= bf.ferm_charge;
    // KSphaseX == 1 everywhere
    phases[(long)(xmap * 2 + parity) * sh + idxh] = bounded ? rebound_2pi(arg) : arg;

    arg = (x - tnx / 2 + 1) * bf.bz / (tnx * tny);
    if (Y == tny) {
        arg -= (z - tnz / 2 + 1) * tny * bf.bx / (tny * tnz);
        arg -= (tt - tnt / 2 + 1) * tny * bf.ey / (tny * tnt);
    }
    arg *= bf.ferm_charge;
    if ((1 - 2 * (x & 1)) == -1) arg += 0.5;
    phases[(long)(ymap * 2 + parity) * sh + idxh] = bounded ? rebound_2pi(arg) : arg;

    arg = (y - tny / 2 + 1) * bf.bx / (tny * tnz);
    if (Z == tnz) {
        arg -= (tt - tnt / 2 + 1) * tnz * bf.ez / (tnz * tnt);
        arg -= (x - tnx / 2 + 1) * tnz * bf.by / (tnz * tnx);
    }
    arg *= bf.ferm_charge;
    if ((1 - 2 * ((x + y) & 1)) == -1) arg += 0.5;
    phases[(long)(zmap * 2 + parity) * sh + idxh] = bounded ? rebound_2pi(arg) : arg;

    arg = (z - tnz / 2 + 1) * bf.ez / (tnz * tnt);
    arg += (y - tny / 2 + 1) * bf.ey / (tny * tnt);
    arg += (x - tnx / 2 + 1) * bf.ex / (tnx * tnt);
    arg *= bf.ferm_charge;
    if ((1 - 2 * ((x + y + z) & 1)) == -1) arg += 0.5;
    arg += chpotphase * 0.5;
    if (T == tnt) arg += 0.5;
    phases[(long)(tmap * 2 + parity) * sh + idxh] = bounded ? rebound_2pi(arg) : arg;
}

__global__ __launch_bounds__(kBlock) void k_d2f_real(const double* __restrict__ d, float* __restrict__ f, long n) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) f[i] = (float)d[i];
}

// ------------------------------------------------------------------ Philox
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += W0; k1 += W1; }
        const uint32_t hi0 = __umulhi(M0, c[0]), lo0 = M0 * c[0];
        const uint32_t hi1 = __umulhi(M1, c[2]), lo1 = M1 * c[2];
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}

enum { kTagGauss = 1, kTagZ2 = 2, kTagCold = 3, kTagExpH = 4, kTagMom = 5 };

__device__ __forceinline__ void draw(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t w[4]) {
    w[0] = a; w[1] = b; w[2] = c; w[3] = d;
    philox4x32_10(w, (uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32));
}

__device__ __forceinline__ double u53(uint32_t x0, uint32_t x1) {
    return ((double)(x0 >> 5) * 67108864.0 + (double)(x1 >> 6)) * (1.0 / 9007199254740992.0);
}

__global__ __launch_bounds__(kBlock) void k_gen_vec3(double2* __restrict__ v, int sizeh, uint64_t seed,
                                                     uint32_t stream, int z2) {
#pragma clang fp contract(off)
    const long i3 = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i3 >= 3L * sizeh) return;
    const int c = (int)(i3 / sizeh), i = (int)(i3 - (long)c * sizeh);
    uint32_t w[4];
    draw(seed, (uint32_t)i, (uint32_t)c, stream, z2 ? kTagZ2 : kTagGauss, w);
    if (z2) {
        v[i3] = make_double2(u53(w[0], w[1]) < 0.5 ? 1.0 : -1.0, 0.0);
    } else {
        // two_double_gauss (random_assignement.c:44-54): radius = sqrt(-log u), E|z|^2 = 1
        const double u1 = u53(w[0], w[1]), u2 = u53(w[2], w[3]);
        const double phi = (2 * 3.14159265358979323846) * u2;
        const double radius = sqrt(-1.0 * log(1.0 - u1));
        v[i3] = make_double2(radius * cos(phi), radius * sin(phi));
    }
}

__global__ __launch_bounds__(kBlock) void k_gen_conf(double2* __restrict__ u, size_t su3_stride, int sizeh, double f,
                                                     uint64_t seed, int kind, int par_temp) {
#pragma clang fp contract(off)
    const long gid = (long)blockIdx.x * kBlock + threadIdx.x;
    if (gid >= 8L * sizeh) return;
    const int link = (int)(gid / sizeh), i = (int)(gid - (long)link * sizeh);
    Su3 m;
    uint32_t w[4];
    if (kind == 0) {
        // generate_random_su3 (random_assignement.c:106-142)
        for (int e = 0; e < 6; e++) {
            draw(seed, (uint32_t)i, (uint32_t)(8 * link + e), 0, kTagCold, w);
            const double a = 1.0 - 2.0 * u53(w[0], w[1]);
            const double b = 1.0 - 2.0 * u53(w[2], w[3]);
            double2 z = make_double2(a * f, b * f);
            if (e == 0 || e == 4) z.x = 1 + z.x;
            m.c[e / 3][e % 3] = z;
        }
        double norm = m.c[0][0].x * m.c[0][0].x + m.c[0][0].y * m.c[0][0].y;
        norm += m.c[0][1].x * m.c[0][1].x + m.c[0][1].y * m.c[0][1].y;
        norm += m.c[0][2].x * m.c[0][2].x + m.c[0][2].y * m.c[0][2].y;
        norm = 1.0 / sqrt(norm);
        for (int c = 0; c < 3; c++) m.c[0][c] = make_double2(m.c[0][c].x * norm, m.c[0][c].y * norm);
        double2 prod = cmul_nc(cconj(m.c[0][0]), m.c[1][0]);
        prod = cadd(prod, cmul_nc(cconj(m.c[0][1]), m.c[1][1]));
        prod = cadd(prod, cmul_nc(cconj(m.c[0][2]), m.c[1][2]));
        for (int c = 0; c < 3; c++) m.c[1][c] = csub(m.c[1][c], cmul_nc(prod, m.c[0][c]));
        norm = m.c[1][0].x * m.c[1][0].x + m.c[1][0].y * m.c[1][0].y;
        norm += m.c[1][1].x * m.c[1][1].x + m.c[1][1].y * m.c[1][1].y;
        norm += m.c[1][2].x * m.c[1][2].x + m.c[1][2].y * m.c[1][2].y;
        norm = 1.0 / sqrt(norm);
        for (int c = 0; c < 3; c++) m.c[1][c] = make_double2(m.c[1][c].x * norm, m.c[1][c].y * norm);
    } else {
        // exp(i beta H): H traceless Hermitian, standard-normal entries
        double n[10];
        for (int k = 0; k < 5; k++) {
            draw(seed, (uint32_t)i, (uint32_t)(8 * link + k), 0, kTagExpH, w);
            const double u1 = u53(w[0], w[1]), u2 = u53(w[2], w[3]);
            const double rad = sqrt(-2.0 * log(1.0 - u1));
            const double phi = (2 * 3.14159265358979323846) * u2;
            n[2 * k] = rad * cos(phi);
            n[2 * k + 1] = rad * sin(phi);
        }
        const double tr3 = (n[0] + n[1] + n[2]) / 3.0;
        const double h00 = n[0] - tr3, h11 = n[1] - tr3;
        Tamat qa;  // QA = -i beta H
        qa.c01 = make_double2(f * n[4], -(f * n[3]));
        qa.c02 = make_double2(f * n[6], -(f * n[5]));
        qa.c12 = make_double2(f * n[8], -(f * n[7]));
        qa.ic00 = -f * h00;
        qa.ic11 = -f * h11;
        ch_exp_minus_qa(qa, m);
    }
    rebuild3row(m);
    double2* U = (double2*)((char*)u + link * su3_stride);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) U[(long)(3 * r + c) * sizeh + i] = m.c[r][c];
    if (par_temp) ((double*)(U + 9L * sizeh))[i] = 1.0;
}

// generate_Momenta_gauss (random_assignement.c:145-185): eight normals of
// variance 1/2 per link (four two_double_gauss pairs), combined into the
// traceless Hermitian momentum the reference builds from them
__global__ __launch_bounds__(kBlock) void k_gen_momenta(void* __restrict__ mom, int sizeh, uint64_t seed,
                                                        uint32_t stream) {
#pragma clang fp contract(off)
    const long gid = (long)blockIdx.x * kBlock + threadIdx.x;
    if (gid >= 8L * sizeh) return;
    const int link = (int)(gid / sizeh), i = (int)(gid - (long)link * sizeh);
    double r[8];
    for (int k = 0; k < 4; k++) {
        uint32_t w[4];
        draw(seed, (uint32_t)i, (uint32_t)(4 * link + k), stream, kTagMom, w);
        const double u1 = u53(w[0], w[1]), u2 = u53(w[2], w[3]);
        const double phi = (2 * 3.14159265358979323846) * u2;
        const double radius = sqrt(-1.0 * log(1.0 - u1));
        r[2 * k] = radius * cos(phi);
        r[2 * k + 1] = radius * sin(phi);
    }
    const double one_by_sqrt3 = 1.0 / sqrt(3.0);
    Thmat h;
    h.rc00 = r[2] + r[7] * one_by_sqrt3;
    h.rc11 = -r[2] + r[7] * one_by_sqrt3;
    h.c01 = make_double2(r[0], -r[1]);
    h.c02 = make_double2(r[3], -r[4]);
    h.c12 = make_double2(r[5], -r[6]);
    th_store(mom, link, sizeh, i, h);
}

// Streaming probes: the HBM ceiling the Dslash is compared with.  KIND 0
// copies a -> b, 1 only reads a, 2 only writes b.  Each block takes chunks of
// U x 256 16-byte elements (U independent 128-bit loads in flight per thread,
// non-temporal), chunks dealt grid-stride.  The read probe folds its loads
// into one value per thread that is stored only if it equals a sentinel the
// data never produce, so the loads cannot be dropped.  n must be a multiple
// of U x 256 (checked by the host).
template <int KIND, int U>
__global__ __launch_bounds__(kBlock) void k_stream(const double2* __restrict__ a, double2* __restrict__ b, long n,
                                                   double* __restrict__ sink) {
    const long nchunk = n / ((long)U * kBlock);
    double acc = 0.0;
    for (long c = blockIdx.x; c < nchunk; c += gridDim.x) {
        const long base = c * U * kBlock + threadIdx.x;
        double2 v[U];
        if (KIND != 2) {
#pragma unroll
            for (int k = 0; k < U; k++) v[k] = ld_nt(a + base + (long)k * kBlock);
        }
#pragma unroll
        for (int k = 0; k < U; k++) {
            if (KIND == 0) st_nt(b + base + (long)k * kBlock, v[k]);
            if (KIND == 1) acc += v[k].x + v[k].y;
            if (KIND == 2) st_nt(b + base + (long)k * kBlock, make_double2((double)c, (double)k));
        }
    }
    if (KIND == 1 && acc == -1.2345e300) sink[threadIdx.x] = acc;
}

static GeomArgs geom_args() {
    const Lattice& L = ctx().lat;
    GeomArgs g;
    for (int i = 0; i < 4; i++) {
        g.nd[i] = L.nd[i];
        g.map[i] = L.map[i];
    }
    g.sizeh = (int)L.sizeh;
    return g;
}

}  // namespace ostp

using namespace ostp;

namespace ostp {
// calc_u1_phases_unb_no2pi on a device array of 8 double_soa; waits
bool u1_phases_unbounded(double* phases, bf_param b, double im_chem_pot, double charge) {
    BfArgs bf{b.ex, b.ey, b.ez, b.bx, b.by, b.bz, im_chem_pot, charge};
    hipLaunchKernelGGL(k_u1_phases, dim3(blocks_for(2 * ctx().lat.sizeh)), dim3(kBlock), 0, ctx().stream, phases,
                       geom_args(), bf, 0);
    return hip_ok(hipGetLastError(), "k_u1_phases") && hip_ok(hipStreamSynchronize(ctx().stream), "k_u1_phases");
}
}  // namespace ostp

extern "C" {

int ostp_compute_nnp_and_nnm(int* nnp, int* nnm) {
    if (!require_init("ostp_compute_nnp_and_nnm")) return 1;
    const Lattice& L = ctx().lat;
    const size_t bytes = (size_t)L.sizeh * 8 * sizeof(int);
    int *dp = nullptr, *dm = nullptr;
    if (!hip_ok(hipMalloc(&dp, bytes), "nn alloc") || !hip_ok(hipMalloc(&dm, bytes), "nn alloc")) return 1;
    const long vol = 2 * L.sizeh;
    hipLaunchKernelGGL(k_nn_tables, dim3(blocks_for(vol)), dim3(kBlock), 0, ctx().stream, dp, dm, geom_args());
    bool ok = hip_ok(hipGetLastError(), "k_nn_tables") && hip_ok(hipStreamSynchronize(ctx().stream), "sync") &&
              ostp_memcpy_dtoh(nnp, dp, bytes) == 0 && ostp_memcpy_dtoh(nnm, dm, bytes) == 0;
    (void)hipFree(dp);
    (void)hipFree(dm);
    return ok ? 0 : 1;
}

void calc_u1_phases(double_soa* phases, bf_param bf_pars, double im_chem_pot, double ferm_charge) {
    if (!require_init("calc_u1_phases")) return;
    const Lattice& L = ctx().lat;
    double* d = (double*)resolve(phases, 8 * L.dsoa_bytes(), "calc_u1_phases");
    if (!d) return;
    BfArgs bf{bf_pars.ex, bf_pars.ey, bf_pars.ez, bf_pars.bx, bf_pars.by, bf_pars.bz, im_chem_pot, ferm_charge};
    hipLaunchKernelGGL(k_u1_phases, dim3(blocks_for(2 * L.sizeh)), dim3(kBlock), 0, ctx().stream, d, geom_args(), bf,
                       1);
    if (!hip_ok(hipGetLastError(), "k_u1_phases") || !hip_ok(hipStreamSynchronize(ctx().stream), "sync")) return;
    // a registered host array is updated too, as the reference computes it on the host
    if (ostp_acc_is_present(phases, 8 * L.dsoa_bytes()) && (void*)d != (void*)phases)
        ostp_memcpy_dtoh(phases, d, 8 * L.dsoa_bytes());
}

void ostp_phases_to_float(const double_soa* phases, float_soa* phases_f) {
    if (!require_init("ostp_phases_to_float")) return;
    const Lattice& L = ctx().lat;
    const double* d = (const double*)resolve(phases, 8 * L.dsoa_bytes(), "ostp_phases_to_float");
    float* f = (float*)resolve(phases_f, 8 * L.fsoa_bytes(), "ostp_phases_to_float");
    if (!d || !f) return;
    const long n = 8 * L.sizeh;
    hipLaunchKernelGGL(k_d2f_real, dim3(blocks_for(n)), dim3(kBlock), 0, ctx().stream, d, f, n);
    if (hip_ok(hipGetLastError(), "k_d2f_real")) hip_ok(hipStreamSynchronize(ctx().stream), "sync");
}

static void gen_vec3(vec3_soa* v, uint64_t seed, uint32_t stream, int z2, const char* nm) {
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    double2* d = (double2*)resolve(v, L.vec3_bytes(), nm);
    if (!d) return;
    hipLaunchKernelGGL(k_gen_vec3, dim3(blocks_for(3 * L.sizeh)), dim3(kBlock), 0, ctx().stream, d, (int)L.sizeh,
                       seed, stream, z2);
    if (hip_ok(hipGetLastError(), nm)) hip_ok(hipStreamSynchronize(ctx().stream), nm);
}

void ostp_generate_vec3_soa_gauss(vec3_soa* v, uint64_t seed, uint32_t stream) {
    gen_vec3(v, seed, stream, 0, "ostp_generate_vec3_soa_gauss");
}
void ostp_generate_vec3_soa_z2noise(vec3_soa* v, uint64_t seed, uint32_t stream) {
    gen_vec3(v, seed, stream, 1, "ostp_generate_vec3_soa_z2noise");
}

static void gen_conf(su3_soa* conf, double f, uint64_t seed, int kind, const char* nm) {
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    double2* d = (double2*)resolve(conf, 8 * L.su3_bytes(), nm);
    if (!d) return;
    hipLaunchKernelGGL(k_gen_conf, dim3(blocks_for(8 * L.sizeh)), dim3(kBlock), 0, ctx().stream, d, L.su3_bytes(),
                       (int)L.sizeh, f, seed, kind, L.par_temp);
    if (hip_ok(hipGetLastError(), nm)) hip_ok(hipStreamSynchronize(ctx().stream), nm);
}

void ostp_generate_Conf_cold(su3_soa* conf, double factor, uint64_t seed) {
    gen_conf(conf, factor, seed, 0, "ostp_generate_Conf_cold");
}
void ostp_generate_Conf_expH(su3_soa* conf, double beta, uint64_t seed) {
    gen_conf(conf, beta, seed, 1, "ostp_generate_Conf_expH");
}

void ostp_generate_Momenta_gauss(thmat_soa* mom8, uint64_t seed, uint32_t stream) {
    const char* nm = "ostp_generate_Momenta_gauss";
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    void* d = resolve(mom8, (size_t)8 * 64 * L.sizeh, nm);
    if (!d) return;
    hipLaunchKernelGGL(k_gen_momenta, dim3(blocks_for(8 * L.sizeh)), dim3(kBlock), 0, ctx().stream, d, (int)L.sizeh,
                       seed, stream);
    if (hip_ok(hipGetLastError(), nm)) hip_ok(hipStreamSynchronize(ctx().stream), nm);
}

double ostp_time_stream(int kind, void* dst, const void* src, size_t bytes, int reps, int grid, int unroll) {
    const char* nm = "ostp_time_stream";
    if (!require_init(nm)) return -1.0;
    if (kind < 0 || kind > 2 || reps <= 0 || grid <= 0) return fail("%s: bad kind/reps/grid", nm), -1.0;
    if (unroll != 1 && unroll != 2 && unroll != 4 && unroll != 8) return fail("%s: unroll must be 1, 2, 4 or 8", nm), -1.0;
    if (bytes % ((size_t)16 * kBlock * unroll) != 0)
        return fail("%s: bytes must be a multiple of %d", nm, 16 * kBlock * unroll), -1.0;
    // the read probe's never-taken store still needs a valid address: the first 256 doubles of src
    double2* b = kind == 1 ? nullptr : (double2*)resolve(dst, bytes, nm);
    const double2* a = kind == 2 ? nullptr : (const double2*)resolve(src, bytes, nm);
    if ((kind != 1 && !b) || (kind != 2 && !a)) return -1.0;
    double* sink = (double*)(kind == 1 ? (void*)a : (void*)b);
    const long n = (long)(bytes / 16);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double us = -1.0;
    if (hip_ok(hipEventCreate(&e0), "event") && hip_ok(hipEventCreate(&e1), "event") &&
        hip_ok(hipEventRecord(e0, ctx().stream), "event record")) {
        for (int r = 0; r < reps; r++) {
#define OSTP_S(K, U) hipLaunchKernelGGL((k_stream<K, U>), dim3(grid), dim3(kBlock), 0, ctx().stream, a, b, n, sink)
#define OSTP_SK(K) if (unroll == 1) OSTP_S(K, 1); else if (unroll == 2) OSTP_S(K, 2); \
                   else if (unroll == 4) OSTP_S(K, 4); else OSTP_S(K, 8)
            if (kind == 0) { OSTP_SK(0); } else if (kind == 1) { OSTP_SK(1); } else { OSTP_SK(2); }
#undef OSTP_SK
#undef OSTP_S
        }
        float ms = 0.f;
        if (hip_ok(hipGetLastError(), "k_stream") && hip_ok(hipEventRecord(e1, ctx().stream), "event record") &&
            hip_ok(hipEventSynchronize(e1), "event sync") && hip_ok(hipEventElapsedTime(&ms, e0, e1), "elapsed"))
            us = 1000.0 * (double)ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return us;
}

}  // extern "C"
