// Fermion force of the staggered RHMC on gfx950 (FP64).
//
//  ker_openacc_compute_fermion_force      OpenAcc/fermion_force_utilities.c:183-201
//  direct_product_of_fermions_into_auxmat                                  :30-93
//  multiply_backfield_times_force                                          :123-153
//  multiply_conf_times_force_and_take_ta_nophase                           :95-121
//  fermion_force_soloopenacc              OpenAcc/fermion_force.c:166-358
//
// The reference applies Doe and accumulates the outer products one shift at
// a time, re-reading and re-writing the 8 gl(3) force arrays per shift.  Here
// all D_oe x_k are formed first and one kernel accumulates sum_k a_k (outer
// products) per link in registers, so the force arrays are read and written
// once per pseudofermion instead of once per shift.
#include "ostp_blas.h"
#include "ostp_su3.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

namespace ostp {

struct FGeom {
    int nd[4];
    int sizeh;
};

__device__ __forceinline__ int nnp_index(int idxh, int parity, int mu, const FGeom& g) {
    const int nd0h = g.nd[0] >> 1;
    int t = idxh / nd0h;
    const int hd0 = idxh - t * nd0h;
    int t2 = t / g.nd[1];
    int d[4];
    d[1] = t - t2 * g.nd[1];
    d[3] = t2 / g.nd[2];
    d[2] = t2 - d[3] * g.nd[2];
    d[0] = 2 * hd0 + ((d[1] + d[2] + d[3] + parity) & 1);
    d[mu] += 1;
    if (d[mu] == g.nd[mu]) d[mu] = 0;
    return (d[0] + g.nd[0] * (d[1] + g.nd[1] * (d[2] + g.nd[2] * d[3]))) >> 1;
}

struct ShiftCoeffs {
    double a[kMaxShifts];
    int order;
};

// aux[2mu+p](x) += sum_k c_k l_k(x+mu) (x) r_k(x)^+ with
//   p = 0 (even x): l = h_k = D_oe x_k (odd field), r = x_k (even), c_k = +a_k
//   p = 1 (odd x) : l = x_k (even field),          r = h_k (odd),  c_k = -a_k
__global__ __launch_bounds__(kBlock) void k_direct_product(double2* __restrict__ aux, size_t su3_stride,
                                                          const double2* __restrict__ xs, const double2* __restrict__ hs,
                                                          long vstride, ShiftCoeffs co, FGeom g) {
    const int gid = xcd_block_id() * kBlock + (int)threadIdx.x;
    const int sh = g.sizeh;
    if (gid >= 8 * sh) return;
    const int dir = gid / sh, idxh = gid - dir * sh;
    const int mu = dir >> 1, parity = dir & 1;
    const int ip = nnp_index(idxh, parity, mu, g);
    Su3 acc;
    su3_zero(acc);
    for (int k = 0; k < co.order; k++) {
        const double2* X = xs + k * vstride;
        const double2* H = hs + k * vstride;
        const double2* Lf = parity == 0 ? H : X;
        const double2* Rf = parity == 0 ? X : H;
        const double c = parity == 0 ? co.a[k] : -co.a[k];
        double2 l[3], r[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            l[a] = Lf[a * sh + ip];
            r[a] = cscale(cconj(Rf[a * sh + idxh]), c);
        }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) acc.c[a][b] = cfma(l[a], r[b], acc.c[a][b]);
    }
    double2* A = (double2*)((char*)aux + dir * su3_stride);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) A[(3 * a + b) * sh + idxh] = cadd(A[(3 * a + b) * sh + idxh], acc.c[a][b]);
}

// The same sums with one thread per half-site and NMU directions: the
// thread's own vector (r) is read once per shift for all NMU links of its
// site instead of once per link, and the neighbours' vectors come from the
// L2 of the XCD that owns the slab (per-link threads read all 2 x order
// vectors 8 times over: 9.7 GB per launch at 48^3x12 and 19 shifts).  Each
// link's sum runs over the shifts in the same order as k_direct_product, so
// the results are bit-identical.  Thread = (parity, direction group, half-site).
template <int NMU>
__global__ __launch_bounds__(kBlock) void k_direct_product_sites(double2* __restrict__ aux, size_t su3_stride,
                                                                const double2* __restrict__ xs,
                                                                const double2* __restrict__ hs, long vstride,
                                                                ShiftCoeffs co, FGeom g) {
    const int gid = xcd_block_id() * kBlock + (int)threadIdx.x;
    const int sh = g.sizeh;
    constexpr int NG = 4 / NMU;
    if (gid >= 2 * NG * sh) return;
    const int pg = gid / sh, idxh = gid - pg * sh;
    const int parity = pg / NG, mu0 = (pg - parity * NG) * NMU;
    int ip[NMU];
#pragma unroll
    for (int m = 0; m < NMU; m++) ip[m] = nnp_index(idxh, parity, mu0 + m, g);
    Su3 acc[NMU];
#pragma unroll
    for (int m = 0; m < NMU; m++) su3_zero(acc[m]);
    for (int k = 0; k < co.order; k++) {
        const double2* X = xs + k * vstride;
        const double2* H = hs + k * vstride;
        const double2* Lf = parity == 0 ? H : X;
        const double2* Rf = parity == 0 ? X : H;
        const double c = parity == 0 ? co.a[k] : -co.a[k];
        double2 r[3];
#pragma unroll
        for (int a = 0; a < 3; a++) r[a] = cscale(cconj(Rf[a * sh + idxh]), c);
#pragma unroll
        for (int m = 0; m < NMU; m++) {
            double2 l[3];
#pragma unroll
            for (int a = 0; a < 3; a++) l[a] = Lf[a * sh + ip[m]];
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) acc[m].c[a][b] = cfma(l[a], r[b], acc[m].c[a][b]);
        }
    }
#pragma unroll
    for (int m = 0; m < NMU; m++) {
        double2* A = (double2*)((char*)aux + (2 * (mu0 + m) + parity) * su3_stride);
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++)
                A[(3 * a + b) * sh + idxh] = cadd(A[(3 * a + b) * sh + idxh], acc[m].c[a][b]);
    }
}

// out[dir] += aux[dir] * e^{i phase[dir]}
__global__ __launch_bounds__(kBlock) void k_backfield_times_force(const double2* __restrict__ aux,
                                                                 double2* __restrict__ out, size_t su3_stride,
                                                                 const double* __restrict__ phases, int sizeh) {
    const int gid = blockIdx.x * kBlock + (int)threadIdx.x;
    if (gid >= 8 * sizeh) return;
    const int dir = gid / sizeh, i = gid - dir * sizeh;
    double s, c;
    phase_sincos(phases[(long)dir * sizeh + i], &s, &c);
    const double2 ph = make_double2(c, s);
    const double2* A = (const double2*)((const char*)aux + dir * su3_stride);
    double2* O = (double2*)((char*)out + dir * su3_stride);
#pragma unroll
    for (int k = 0; k < 9; k++) O[k * sizeh + i] = cfma(A[k * sizeh + i], ph, O[k * sizeh + i]);
}

// ipdot[dir] -= TA(U[dir] * aux[dir])
__global__ __launch_bounds__(kBlock) void k_conf_times_force_ta(const double2* __restrict__ u,
                                                               const double2* __restrict__ aux, size_t su3_stride,
                                                               void* __restrict__ ipdot, int sizeh) {
    const int gid = blockIdx.x * kBlock + (int)threadIdx.x;
    if (gid >= 8 * sizeh) return;
    const int dir = gid / sizeh, i = gid - dir * sizeh;
    const double2* U = (const double2*)((const char*)u + dir * su3_stride);
    const double2* A = (const double2*)((const char*)aux + dir * su3_stride);
    Su3 m, a, x;
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) m.c[r][c] = U[(3 * r + c) * sizeh + i];
    rebuild3row(m);
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) a.c[r][c] = A[(3 * r + c) * sizeh + i];
    su3_mul(m, a, x);
    Tamat t;
    ta_part(x, 1.0, t);
    char* b = (char*)ipdot + (size_t)dir * 64 * sizeh;
    double2* pc = (double2*)b;
    double* pd = (double*)(b + (size_t)48 * sizeh);
    pc[i] = csub(pc[i], t.c01);
    pc[sizeh + i] = csub(pc[sizeh + i], t.c02);
    pc[2 * sizeh + i] = csub(pc[2 * sizeh + i], t.c12);
    pd[i] -= t.ic00;
    pd[sizeh + i] -= t.ic11;
}

static FGeom fgeom() {
    const Lattice& L = ctx().lat;
    FGeom g;
    for (int i = 0; i < 4; i++) g.nd[i] = L.nd[i];
    g.sizeh = (int)L.sizeh;
    return g;
}
static int nblk8() { return blocks_for(8 * ctx().lat.sizeh); }

static int g_dp_kernel = 4;  // links of a site per thread in the outer-product kernel (0: one thread per link)

bool launch_direct_product(void* aux, const void* xs, const void* hs, const double* a, int order) {
    ShiftCoeffs co;
    co.order = order;
    for (int k = 0; k < kMaxShifts; k++) co.a[k] = k < order ? a[k] : 0.0;
    const long sh = ctx().lat.sizeh;
    const size_t st = ctx().lat.su3_bytes();
    if (g_dp_kernel == 4)
        hipLaunchKernelGGL((k_direct_product_sites<4>), dim3(blocks_for(2 * sh)), dim3(kBlock), 0, ctx().stream,
                           (double2*)aux, st, (const double2*)xs, (const double2*)hs, 3 * sh, co, fgeom());
    else if (g_dp_kernel == 2)
        hipLaunchKernelGGL((k_direct_product_sites<2>), dim3(blocks_for(4 * sh)), dim3(kBlock), 0, ctx().stream,
                           (double2*)aux, st, (const double2*)xs, (const double2*)hs, 3 * sh, co, fgeom());
    else
        hipLaunchKernelGGL(k_direct_product, dim3(nblk8()), dim3(kBlock), 0, ctx().stream, (double2*)aux, st,
                           (const double2*)xs, (const double2*)hs, 3 * sh, co, fgeom());
    return hip_ok(hipGetLastError(), "k_direct_product");
}

bool launch_backfield_times_force(const void* aux, void* out, const void* phases) {
    hipLaunchKernelGGL(k_backfield_times_force, dim3(nblk8()), dim3(kBlock), 0, ctx().stream, (const double2*)aux,
                       (double2*)out, ctx().lat.su3_bytes(), (const double*)phases, (int)ctx().lat.sizeh);
    return hip_ok(hipGetLastError(), "k_backfield_times_force");
}

bool launch_conf_times_force_ta(const void* u, const void* aux, void* ipdot) {
    hipLaunchKernelGGL(k_conf_times_force_ta, dim3(nblk8()), dim3(kBlock), 0, ctx().stream, (const double2*)u,
                       (const double2*)aux, ctx().lat.su3_bytes(), ipdot, (int)ctx().lat.sizeh);
    return hip_ok(hipGetLastError(), "k_conf_times_force_ta");
}

bool launch_stout(const void* u, void* uout, double rho, double c0, void* stap, void* qa, void* expo);
bool launch_sigma_chain(void* sigma, void* lam, void* qa, const void* u, double rho, double c0);
double stout_c_zero();

// library-owned scratch (the reference's globals aux_th, aux_ta, conf_acc_f and
// the per-shift D_oe x_k fields)
struct Scratch {
    void* p = nullptr;
    size_t n = 0;
    void* get(size_t need) {
        if (n >= need) return p;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (!hip_ok(hipMalloc(&p, need), "force scratch")) return nullptr;
        n = need;
        return p;
    }
};
static Scratch g_hs, g_th, g_ta, g_conf_f, g_xs_all;

// the rational solves of several sources on one operator, together (solvers.hip)
bool multishift_fp64_many(const void* u, ferm_param* const* pars, const RationalApprox* const* ras,
                          const void* const* in, void* const* out, int n, double res, int max_cg,
                          int convergence_importance);

// sum over shifts of the outer products for one pseudofermion
bool compute_fermion_force(const void* u, void* aux, const void* xs, const void* phases, const RationalApprox* ra) {
    const Lattice& L = ctx().lat;
    const int order = ra->approx_order;
    if (order <= 0) return true;
    void* hs = g_hs.get((size_t)order * L.vec3_bytes());
    if (!hs) return false;
    // all D_oe x_k, the shifts' solutions in chunks through shared-link hops
    if (!launch_doe_multi(u, phases, xs, hs, order)) return false;
    return launch_direct_product(aux, xs, hs, ra->RA_a, order);
}

}  // namespace ostp

using namespace ostp;

extern "C" {

extern double gl_stout_rho;
int ostp_stout_steps(void);

int ostp_set_direct_product_kernel(int nmu) {
    if (nmu != 0 && nmu != 2 && nmu != 4) {
        fail("ostp_set_direct_product_kernel: %d directions per thread (0, 2 or 4)", nmu);
        return -1;
    }
    const int prev = g_dp_kernel;
    g_dp_kernel = nmu;
    return prev;
}

void direct_product_of_fermions_into_auxmat(const vec3_soa* loc_s, const vec3_soa* loc_h, su3_soa* aux_u,
                                            const RationalApprox* approx, int iter) {
    const char* nm = "direct_product_of_fermions_into_auxmat";
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    const void* ds = resolve(loc_s, L.vec3_bytes(), nm);
    const void* dh = resolve(loc_h, L.vec3_bytes(), nm);
    void* da = resolve(aux_u, 8 * L.su3_bytes(), nm);
    if (!ds || !dh || !da || !approx) return;
    if (iter < 0 || iter >= approx->approx_order) { fail("%s: shift %d out of range", nm, iter); return; }
    double a = approx->RA_a[iter];
    if (launch_direct_product(da, ds, dh, &a, 1)) hip_ok(hipStreamSynchronize(ctx().stream), nm);
}

void multiply_backfield_times_force(ferm_param* tpars, const su3_soa* auxmat, su3_soa* pseudo_ipdot) {
    const char* nm = "multiply_backfield_times_force";
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    const void* dph = resolve(tpars ? tpars->phases : nullptr, 8 * L.dsoa_bytes(), nm);
    const void* da = resolve(auxmat, 8 * L.su3_bytes(), nm);
    void* dout = resolve(pseudo_ipdot, 8 * L.su3_bytes(), nm);
    if (!dph || !da || !dout) return;
    if (launch_backfield_times_force(da, dout, dph)) hip_ok(hipStreamSynchronize(ctx().stream), nm);
}

void multiply_conf_times_force_and_take_ta_nophase(const su3_soa* u, const su3_soa* auxmat, tamat_soa* ipdot) {
    const char* nm = "multiply_conf_times_force_and_take_ta_nophase";
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    const void* du = resolve(u, 8 * L.su3_bytes(), nm);
    const void* da = resolve(auxmat, 8 * L.su3_bytes(), nm);
    void* dp = resolve(ipdot, (size_t)8 * 64 * L.sizeh, nm);
    if (!du || !da || !dp) return;
    if (launch_conf_times_force_ta(du, da, dp)) hip_ok(hipStreamSynchronize(ctx().stream), nm);
}

void ker_openacc_compute_fermion_force(const su3_soa* u, su3_soa* aux_u, const vec3_soa* in_shiftmulti,
                                       vec3_soa* loc_s, vec3_soa* loc_h, ferm_param* tpars) {
    const char* nm = "ker_openacc_compute_fermion_force";
    (void)loc_s;
    (void)loc_h;
    if (!require_init(nm)) return;
    if (!tpars) { fail("%s: null ferm_param", nm); return; }
    const Lattice& L = ctx().lat;
    const int order = tpars->approx_md.approx_order;
    const void* du = resolve(u, 8 * L.su3_bytes(), nm);
    void* da = resolve(aux_u, 8 * L.su3_bytes(), nm);
    const void* dx = resolve(in_shiftmulti, (size_t)(order > 0 ? order : 1) * L.vec3_bytes(), nm);
    const void* dph = resolve(tpars->phases, 8 * L.dsoa_bytes(), nm);
    if (!du || !da || !dx || !dph) return;
    if (compute_fermion_force(du, da, dx, dph, &tpars->approx_md)) hip_ok(hipStreamSynchronize(ctx().stream), nm);
}

int inverter_multishift_wrapper(inverter_package ip, ferm_param* pars, RationalApprox* approx, vec3_soa* out,
                                const vec3_soa* in, double res, int max_cg, int convergence_importance);
void ostp_convert_double_to_float_gauge(const su3_soa* d_var, su3_soa_f* f_var);

// fermion_force_soloopenacc (fermion_force.c:166-358) for the build with
// STOUT_FERMIONS; the inverter-recycling branch (md_parameters.recycleInvsForce)
// is not part of this entry point.
void fermion_force_soloopenacc(su3_soa* tconf_acc, su3_soa* tstout_conf_acc_arr, su3_soa* gl3_aux,
                               tamat_soa* tipdot_acc, ferm_param* tfermion_parameters, int tNDiffFlavs,
                               const vec3_soa* ferm_in_acc, double res, su3_soa* taux_conf_acc,
                               vec3_soa* tferm_shiftmulti_acc, inverter_package ipt, const int max_cg) {
    const char* nm = "fermion_force_soloopenacc";
    if (!require_init(nm)) return;
    ForceScope in_force;  // inverter_multishift_wrapper's force-only precision mode applies
    const Lattice& L = ctx().lat;
    const size_t gb = 8 * L.su3_bytes(), tb = (size_t)8 * 64 * L.sizeh;
    const int steps = ostp_stout_steps();
    void* dconf = resolve(tconf_acc, gb, nm);
    char* darr = steps > 0 ? (char*)resolve(tstout_conf_acc_arr, steps * gb, nm) : nullptr;
    void* dgl3 = resolve(gl3_aux, gb, nm);
    void* dip = resolve(tipdot_acc, tb, nm);
    void* dtaux = resolve(taux_conf_acc, gb, nm);
    if (!dconf || (steps > 0 && !darr) || !dgl3 || !dip || !dtaux || !tfermion_parameters) return;
    hipStream_t s = ctx().stream;
    if (steps > 0) {
        stout_wrapper(tconf_acc, tstout_conf_acc_arr, 0);
    }
    const void* conf_to_use = steps > 0 ? (const void*)(darr + (steps - 1) * gb) : dconf;
    if (!hip_ok(hipMemsetAsync(dgl3, 0, gb, s), nm) || !hip_ok(hipMemsetAsync(dip, 0, tb, s), nm)) return;
    ipt.u = (const su3_soa*)conf_to_use;
    if (inverter_tricks.singlePInvAccelMultiInv) {
        void* cf = g_conf_f.get(8 * L.su3f_bytes());
        if (!cf) return;
        ostp_convert_double_to_float_gauge((const su3_soa*)conf_to_use, (su3_soa_f*)cf);
        ipt.u_f = (const su3_soa_f*)cf;
    }
    // every (flavour, pseudofermion) source's rational solve first, all together
    // when they share the operator (same phases; FP64 solves), then the forces
    bool together = false;
    {
        const unsigned long e0 = error_count();
        std::vector<ferm_param*> fps;
        std::vector<const RationalApprox*> ras;
        std::vector<const void*> ins;
        std::vector<void*> outs;
        int maxo = 1;
        for (int iflav = 0; iflav < tNDiffFlavs; iflav++) {
            ferm_param* fp = &tfermion_parameters[iflav];
            maxo = std::max(maxo, fp->approx_md.approx_order);
            for (int ips = 0; ips < fp->number_of_ps; ips++) {
                fps.push_back(fp);
                ras.push_back(&fp->approx_md);
                ins.push_back(resolve((const char*)ferm_in_acc + (size_t)(fp->index_of_the_first_ps + ips) *
                                                                     L.vec3_bytes(),
                                      L.vec3_bytes(), nm));
            }
        }
        const int n = (int)fps.size();
        char* xs_all = n >= 2 ? (char*)g_xs_all.get((size_t)n * maxo * L.vec3_bytes()) : nullptr;
        if (n >= 2 && !xs_all) {
            (void)hipGetLastError();
            forgive_errors(e0);
        }
        if (xs_all) {
            for (int k = 0; k < n; k++) {
                if (!ins[k]) return;
                outs.push_back(xs_all + (size_t)k * maxo * L.vec3_bytes());
            }
            together = multishift_fp64_many(conf_to_use, fps.data(), ras.data(), ins.data(), outs.data(), n, res,
                                            max_cg, CONVERGENCE_NONCRITICAL);
            if (!together && error_count() != e0) return;
        }
        if (together) {
            int k = 0;
            for (int iflav = 0; iflav < tNDiffFlavs; iflav++) {
                ferm_param* fp = &tfermion_parameters[iflav];
                if (!hip_ok(hipMemsetAsync(dtaux, 0, gb, s), nm)) return;
                const void* dph = resolve(fp->phases, 8 * L.dsoa_bytes(), nm);
                if (!dph) return;
                for (int ips = 0; ips < fp->number_of_ps; ips++, k++)
                    if (!compute_fermion_force(conf_to_use, dtaux, outs[k], dph, &fp->approx_md)) return;
                if (!launch_backfield_times_force(dtaux, dgl3, dph)) return;
            }
            // the shifted solutions of the last source stay in tferm_shiftmulti_acc, as after the loop below
            const int order = fps[n - 1]->approx_md.approx_order;
            void* dxs = resolve(tferm_shiftmulti_acc, (size_t)(order > 0 ? order : 1) * L.vec3_bytes(), nm);
            if (!dxs || (order > 0 && !hip_ok(hipMemcpyAsync(dxs, outs[n - 1], (size_t)order * L.vec3_bytes(),
                                                             hipMemcpyDeviceToDevice, s),
                                              nm)))
                return;
        }
    }
    for (int iflav = 0; !together && iflav < tNDiffFlavs; iflav++) {
        ferm_param* fp = &tfermion_parameters[iflav];
        if (!hip_ok(hipMemsetAsync(dtaux, 0, gb, s), nm)) return;
        const int ifps = fp->index_of_the_first_ps;
        const void* dph = resolve(fp->phases, 8 * L.dsoa_bytes(), nm);
        const int order = fp->approx_md.approx_order;
        const void* dxs = resolve(tferm_shiftmulti_acc, (size_t)(order > 0 ? order : 1) * L.vec3_bytes(), nm);
        if (!dph || !dxs) return;
        for (int ips = 0; ips < fp->number_of_ps; ips++) {
            const vec3_soa* in = (const vec3_soa*)((const char*)ferm_in_acc + (size_t)(ifps + ips) * L.vec3_bytes());
            inverter_multishift_wrapper(ipt, fp, &fp->approx_md, tferm_shiftmulti_acc, in, res, max_cg,
                                        CONVERGENCE_NONCRITICAL);
            if (!compute_fermion_force(conf_to_use, dtaux, dxs, dph, &fp->approx_md)) return;
        }
        if (!launch_backfield_times_force(dtaux, dgl3, dph)) return;
    }
    if (steps > 0) {
        void* th = g_th.get(tb);
        void* ta = g_ta.get(tb);
        if (!th || !ta) return;
        for (int level = steps; level > 1; level--)
            if (!launch_sigma_chain(dgl3, th, ta, darr + (level - 2) * gb, gl_stout_rho, stout_c_zero())) return;
        if (!launch_sigma_chain(dgl3, th, ta, dconf, gl_stout_rho, stout_c_zero())) return;
    }
    if (launch_conf_times_force_ta(dconf, dgl3, dip)) hip_ok(hipStreamSynchronize(s), nm);
}

}  // extern "C"
