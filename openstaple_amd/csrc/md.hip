// Molecular-dynamics pieces of the RHMC update:
//
//  mom_sum_mult                    OpenAcc/su3_utilities.c:241-270, su3_utilities.h:1383-1394
//  mom_exp_times_conf_soloopenacc  OpenAcc/su3_utilities.c:272-307, su3_utilities.h:1434-1537
//  calc_momenta_action             OpenAcc/su3_measurements.c:138-168, su3_utilities.h:1295-1307
//  invert_momenta                  OpenAcc/su3_measurements.c:170-189
//  initialize_md_global_variables  OpenAcc/md_parameters.c:11-52
//  multistep_2MN_gauge             OpenAcc/md_integrator.c:256-376 (sub-step and loop)
//  multistep_2MN_SOLOOPENACC       OpenAcc/md_integrator.c:502-740
//
// The gauge sub-integrator evaluates each gauge force together with its
// momentum step in one kernel (staples, TA(U S) and P -= i dt TA(U S) per
// link, in registers), so a gauge sub-step is two launches: force + momenta,
// then U <- exp(i dt P) U.  The momenta are one 64-byte record per link
// (thmat_soa: three complex planes, two real planes), the links the
// reference's su3_soa planes; every kernel is one thread per link.
#include "ostp_blas.h"
#include "ostp_su3.h"

#include <cmath>
#include <cstdio>

extern "C" {
// md_parameters.c:8-9,54-55, action.c:6: zero-initialised like the
// reference's globals; a host that defines them takes precedence.
double deltas_Omelyan[7];
float deltas_Omelyan_f[7];
md_param md_parameters;
action_param act_params;
int nMdInversionPerformed;
}

static_assert(sizeof(md_param) == 64, "md_param layout (md_parameters.h:6-19)");
static_assert(offsetof(md_param, t) == 8 && offsetof(md_param, residue_md) == 40 &&
                  offsetof(md_param, extrapolateInvsForce) == 56,
              "md_param layout (md_parameters.h:6-19)");
static_assert(offsetof(action_param, topo_action) == 24 && offsetof(action_param, topo_file_path) == 48 &&
                  sizeof(action_param) == 80,
              "action_param layout (action.h:6-18)");

namespace ostp {

// P += fact * ipdot in the reference's convention (see mom_step)
__global__ __launch_bounds__(kBlock) void k_mom_sum_mult(void* __restrict__ mom, const void* __restrict__ ipdot,
                                                         int sizeh, double fact) {
    const int gid = blockIdx.x * kBlock + (int)threadIdx.x;
    if (gid >= 8 * sizeh) return;
    const int link = gid / sizeh, i = gid - link * sizeh;
    Tamat q;
    ta_load(ipdot, link, sizeh, i, q);
    mom_step(mom, link, sizeh, i, q, fact);
}

// U(rows 0,1) <- rows 0,1 of exp(i delta P) U, then re-unitarised (row 0
// normalised, row 1 orthogonalised to it and normalised); row 2 is not
// written, as in the reference
__global__ __launch_bounds__(kBlock) void k_mom_exp(double2* __restrict__ conf, size_t su3_stride,
                                                    const void* __restrict__ mom, int sizeh, double delta) {
    const int gid = blockIdx.x * kBlock + (int)threadIdx.x;
    if (gid >= 8 * sizeh) return;
    const int link = gid / sizeh, i = gid - link * sizeh;
    Thmat h;
    th_load(mom, link, sizeh, i, h);
    // QA = -i delta P in the tamat form: the reference's -P*delta*1.0I
    Tamat q;
    q.ic00 = -h.rc00 * delta;
    q.ic11 = -h.rc11 * delta;
    q.c01 = make_double2(h.c01.y * delta, -(h.c01.x * delta));
    q.c02 = make_double2(h.c02.y * delta, -(h.c02.x * delta));
    q.c12 = make_double2(h.c12.y * delta, -(h.c12.x * delta));
    Su3 E, U;
    ch_exp_minus_qa(q, E);
    double2* Ul = (double2*)((char*)conf + (size_t)link * su3_stride);
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) U.c[r][c] = Ul[(size_t)(3 * r + c) * sizeh + i];
    rebuild3row(U);
    double2 o[2][3];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double2 s = cmul(E.c[r][0], U.c[0][c]);
            s = cfma(E.c[r][1], U.c[1][c], s);
            o[r][c] = cfma(E.c[r][2], U.c[2][c], s);
        }
    double n = 1.0 / sqrt(cnorm2(o[0][0]) + cnorm2(o[0][1]) + cnorm2(o[0][2]));
#pragma unroll
    for (int c = 0; c < 3; c++) o[0][c] = cscale(o[0][c], n);
    double2 sp = cmul(cconj(o[0][0]), o[1][0]);
    sp = cfma(cconj(o[0][1]), o[1][1], sp);
    sp = cfma(cconj(o[0][2]), o[1][2], sp);
#pragma unroll
    for (int c = 0; c < 3; c++) o[1][c] = csub(o[1][c], cmul(sp, o[0][c]));
    n = 1.0 / sqrt(cnorm2(o[1][0]) + cnorm2(o[1][1]) + cnorm2(o[1][2]));
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) Ul[(size_t)(3 * r + c) * sizeh + i] = r == 0 ? o[0][c] : cscale(o[1][c], n);
}

// sum over the links [first, first + nlinks) of half_tr_thmat_squared = Tr(P^2)/2
__global__ __launch_bounds__(kBlock) void k_mom_action(const void* __restrict__ mom, int first, int nlinks,
                                                       int sizeh, SolverCtl ctl) {
    const int gid = blockIdx.x * kBlock + (int)threadIdx.x;
    double v = 0.0;
    if (gid < nlinks * sizeh) {
        const int l = gid / sizeh, i = gid - l * sizeh;
        Thmat h;
        th_load(mom, first + l, sizeh, i, h);
        v = h.rc00 * h.rc00 + h.rc11 * h.rc11 + h.rc00 * h.rc11 + cnorm2(h.c01) + cnorm2(h.c02) + cnorm2(h.c12);
    }
    block_partial<1>(v, 0.0, ctl);
}

__global__ __launch_bounds__(kBlock) void k_negate(double* __restrict__ a, long n) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = -a[i];
}

static size_t thmat8_bytes() { return (size_t)8 * 64 * ctx().lat.sizeh; }
static int nblk8() { return blocks_for(8 * ctx().lat.sizeh); }

bool launch_mom_sum_mult(void* mom, const void* ipdot, double fact) {
    hipLaunchKernelGGL(k_mom_sum_mult, dim3(nblk8()), dim3(kBlock), 0, ctx().stream, mom, ipdot,
                       (int)ctx().lat.sizeh, fact);
    return hip_ok(hipGetLastError(), "k_mom_sum_mult");
}

bool launch_mom_exp(void* conf, const void* mom, double delta) {
    hipLaunchKernelGGL(k_mom_exp, dim3(nblk8()), dim3(kBlock), 0, ctx().stream, (double2*)conf,
                       ctx().lat.su3_bytes(), mom, (int)ctx().lat.sizeh, delta);
    return hip_ok(hipGetLastError(), "k_mom_exp");
}

// momenta action of links [first, first + nlinks), summed on the device in a
// fixed order and returned on the host
bool momenta_action(const void* mom, int first, int nlinks, double* out) {
    const int sizeh = (int)ctx().lat.sizeh;
    const int nb = blocks_for((long)nlinks * sizeh);
    if (nb > 2 * ctx().max_blocks) return fail("momenta action: %d partial sums exceed the scratch", nb);
    SolverCtl c = plain_ctl();
    hipLaunchKernelGGL(k_mom_action, dim3(nb), dim3(kBlock), 0, ctx().stream, mom, first, nlinks, sizeh, c);
    if (!hip_ok(hipGetLastError(), "k_mom_action") || !launch_finalize(kFinStore, c, nb)) return false;
    double v = 0.0;
    if (!hip_ok(hipMemcpyAsync(&v, &ctx().state->sum, sizeof(double), hipMemcpyDeviceToHost, ctx().stream),
                "momenta action readback") ||
        !hip_ok(hipStreamSynchronize(ctx().stream), "momenta action readback"))
        return false;
    *out = v;
    return true;
}

bool launch_invert_momenta(void* mom) {
    const long n = (long)thmat8_bytes() / 8;
    hipLaunchKernelGGL(k_negate, dim3(blocks_for(n)), dim3(kBlock), 0, ctx().stream, (double*)mom, n);
    return hip_ok(hipGetLastError(), "k_negate");
}

// one gauge sub-step of multistep_2MN_gauge_bloc (md_integrator.c:256-345):
// P -= i deltas[idx] TA(U S) fused with the force, then U <- exp(i deltas[4] P) U
static bool gauge_bloc(void* u, void* ipdot, void* mom, int idx) {
    return launch_gauge_force(u, nullptr, ipdot, mom, deltas_Omelyan[idx]) && launch_mom_exp(u, mom, deltas_Omelyan[4]);
}

// multistep_2MN_gauge (md_integrator.c:347-376) on device pointers
bool md_gauge(void* u, void* stap, void* ipdot, void* mom) {
    const int gs = md_parameters.gauge_scale;
    if (!gauge_bloc(u, ipdot, mom, 3)) return false;
    for (int md = 1; md < gs; md++)
        if (!gauge_bloc(u, ipdot, mom, 5) || !gauge_bloc(u, ipdot, mom, 6)) return false;
    if (!gauge_bloc(u, ipdot, mom, 5)) return false;
    // the last force also leaves its staples, as calc_ipdot_gauge_soloopenacc does
    return launch_gauge_force(u, stap, ipdot, mom, deltas_Omelyan[3]);
}

}  // namespace ostp

using namespace ostp;

static bool sync(const char* nm) { return hip_ok(hipStreamSynchronize(ctx().stream), nm); }

extern "C" {

void mom_sum_mult(thmat_soa* mom, const tamat_soa* ipdot, const double* factor, int id_factor) {
    const char* nm = "mom_sum_mult";
    if (!require_init(nm)) return;
    void* dm = resolve(mom, thmat8_bytes(), nm);
    const void* di = resolve(ipdot, thmat8_bytes(), nm);
    if (!dm || !di) return;
    if (!factor) {
        fail("%s: null factor array", nm);
        return;
    }
    if (launch_mom_sum_mult(dm, di, factor[id_factor])) sync(nm);
}

void mom_exp_times_conf_soloopenacc(su3_soa* tconf_acc, const thmat_soa* mom, const double* factor, int id_factor) {
    const char* nm = "mom_exp_times_conf_soloopenacc";
    if (!require_init(nm)) return;
    void* du = resolve(tconf_acc, 8 * ctx().lat.su3_bytes(), nm);
    const void* dm = resolve(mom, thmat8_bytes(), nm);
    if (!du || !dm) return;
    if (!factor) {
        fail("%s: null factor array", nm);
        return;
    }
    if (launch_mom_exp(du, dm, factor[id_factor])) sync(nm);
}

// tr_local is the reference's work array; the sum is formed on the device
// without it
double calc_momenta_action(const thmat_soa* mom, double_soa* tr_local, const int mu) {
    (void)tr_local;
    const char* nm = "calc_momenta_action";
    if (!require_init(nm)) return NAN;
    const void* dm = resolve(mom, thmat8_bytes(), nm);
    if (!dm) return NAN;
    if (mu < 0 || mu > 7) {
        fail("%s: link array %d out of range 0..7", nm, mu);
        return NAN;
    }
    double v = NAN;
    return momenta_action(dm, mu, 1, &v) ? v : NAN;
}

void invert_momenta(thmat_soa* mom) {
    const char* nm = "invert_momenta";
    if (!require_init(nm)) return;
    void* dm = resolve(mom, thmat8_bytes(), nm);
    if (dm && launch_invert_momenta(dm)) sync(nm);
}

// md_parameters.c:11-52, same arithmetic
void initialize_md_global_variables(md_param p) {
    const int no_md = p.no_md, gauge_scale = p.gauge_scale;
    const double t = p.t;
    const double lambda = 0.1931833275037836;  // Omelyan et al.
    const double beta_by_three = act_params.beta * 0.33333333333333333333333;
    deltas_Omelyan[0] = -t / no_md * lambda;
    deltas_Omelyan[1] = -t / no_md * (1.0 - 2.0 * lambda);
    deltas_Omelyan[2] = -t / no_md * 2.0 * lambda;
    deltas_Omelyan[3] = -t / (2 * no_md * gauge_scale) * lambda * beta_by_three;
    deltas_Omelyan[4] = 0.5 * t / (2 * no_md * gauge_scale);
    deltas_Omelyan[5] = -t / (2 * no_md * gauge_scale) * (1.0 - 2.0 * lambda) * beta_by_three;
    deltas_Omelyan[6] = -t / (2 * no_md * gauge_scale) * 2.0 * lambda * beta_by_three;
    for (int k = 0; k < 7; k++) deltas_Omelyan_f[k] = (float)deltas_Omelyan[k];
}

// local_staples receives the staples of the last force evaluation only
void multistep_2MN_gauge(su3_soa* tconf_acc, su3_soa* local_staples, tamat_soa* tipdot, thmat_soa* tmomenta) {
    const char* nm = "multistep_2MN_gauge";
    if (!require_init(nm)) return;
    const size_t gb = 8 * ctx().lat.su3_bytes();
    void* du = resolve(tconf_acc, gb, nm);
    void* ds = resolve(local_staples, gb, nm);
    void* di = resolve(tipdot, thmat8_bytes(), nm);
    void* dm = resolve(tmomenta, thmat8_bytes(), nm);
    if (!du || !ds || !di || !dm) return;
    if (md_parameters.gauge_scale < 1) {
        fail("%s: gauge_scale %d < 1", nm, md_parameters.gauge_scale);
        return;
    }
    if (verbosity_lv > 1) printf("Performing Gauge substeps\n");
    if (md_gauge(du, ds, di, dm)) sync(nm);
}

// multistep_2MN_SOLOOPENACC for the build with STOUT_FERMIONS: fermion force
// at lambda, 1-2 lambda, 2 lambda of each outer step, gauge sub-integrator
// between them.  The topological force (act_params.topo_action) and the
// force-inversion extrapolation are not part of this entry point.
void multistep_2MN_SOLOOPENACC(tamat_soa* tipdot_acc, su3_soa* tconf_acc, su3_soa* tstout_conf_acc_arr,
                               su3_soa* tauxbis_conf_acc, su3_soa* taux_conf_acc, ferm_param* tfermions_parameters,
                               int tNDiffFlavs, vec3_soa* ferm_in_acc, vec3_soa* tferm_shiftmulti_acc,
                               inverter_package ip, thmat_soa* tmomenta, dcomplex_soa* local_sums, double res,
                               const int max_cg) {
    (void)local_sums;
    const char* nm = "multistep_2MN_SOLOOPENACC";
    if (!require_init(nm)) return;
    if (act_params.topo_action == 1) {
        fail("%s: the topological force is not supported", nm);
        return;
    }
    if (md_parameters.extrapolateInvsForce) {
        fail("%s: extrapolateInvsForce is not implemented (the reference stops there too)", nm);
        return;
    }
    if (md_parameters.no_md < 1 || md_parameters.gauge_scale < 1) {
        fail("%s: no_md %d / gauge_scale %d must be >= 1", nm, md_parameters.no_md, md_parameters.gauge_scale);
        return;
    }
    const Lattice& L = ctx().lat;
    const size_t gb = 8 * L.su3_bytes();
    void* du = resolve(tconf_acc, gb, nm);
    void* dt = resolve(taux_conf_acc, gb, nm);
    void* di = resolve(tipdot_acc, thmat8_bytes(), nm);
    void* dm = resolve(tmomenta, thmat8_bytes(), nm);
    if (!du || !dt || !di || !dm || !tfermions_parameters) return;
    int max_order = 1;
    for (int f = 0; f < tNDiffFlavs; f++)
        if (tfermions_parameters[f].approx_md.approx_order > max_order)
            max_order = tfermions_parameters[f].approx_md.approx_order;
    void* dx = resolve(tferm_shiftmulti_acc, (size_t)max_order * L.vec3_bytes(), nm);
    if (!dx || !hip_ok(hipMemsetAsync(dx, 0, (size_t)max_order * L.vec3_bytes(), ctx().stream), nm)) return;
    nMdInversionPerformed = 0;

    // an error recorded inside the force (non-aborting mode) ends the trajectory
    const unsigned long e0 = error_count();
    auto fermion_step = [&](int idx) {
        fermion_force_soloopenacc(tconf_acc, tstout_conf_acc_arr, tauxbis_conf_acc, tipdot_acc, tfermions_parameters,
                                  tNDiffFlavs, ferm_in_acc, res, taux_conf_acc, tferm_shiftmulti_acc, ip, max_cg);
        if (error_count() != e0) return false;
        return launch_mom_sum_mult(dm, di, deltas_Omelyan[idx]);
    };
    auto gauge = [&]() { return md_gauge(du, dt, di, dm); };

    if (!fermion_step(0)) return;
    for (int md = 1; md < md_parameters.no_md; md++) {
        if (verbosity_lv > 1) printf("RUNNING MD STEP %d OF %d...\n", md, md_parameters.no_md);
        if (!gauge() || !fermion_step(1) || !gauge() || !fermion_step(2)) return;
    }
    if (verbosity_lv > 1) printf("RUNNING LAST MD STEP OF %d...\n", md_parameters.no_md);
    if (!gauge() || !fermion_step(1) || !gauge() || !fermion_step(0)) return;
    sync(nm);
}

}  // extern "C"
