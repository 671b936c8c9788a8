/*
 * Test driver around the reference's own FP64 single-rank sources -- TEST
 * INFRASTRUCTURE ONLY.  Built by oracle/ref.mk from the files where they lie
 * under the reference tree (nothing is copied); output goes to oracle/_ref/.
 *
 * The reference's host program normally owns the globals below (its input
 * parser and allocator, which need MPI / generated FP32 headers and so do not
 * build here); this driver plays that role with plain values.  It exposes
 * small C functions that tests/ call through ctypes to cross-check the oracle.
 */
#include <stdlib.h>
#include <string.h>

#include "OpenAcc/geometry.h"
#include "OpenAcc/struct_c_def.h"
#include "OpenAcc/backfield.h"
#include "OpenAcc/action.h"
#include "Include/debug.h"
#include "OpenAcc/alloc_vars.h"
#include "OpenAcc/fermionic_utilities.h"
#include "RationalApprox/rationalapprox.h"
#include "OpenAcc/plaquettes.h"
#include "OpenAcc/rectangles.h"
#include "OpenAcc/su3_measurements.h"
#include "OpenAcc/su3_utilities.h"
#include "OpenAcc/md_parameters.h"
#include "Meas/baryon_number_utilities.h"

int verbosity_lv = 0;
action_param act_params;
debug_settings_t debug_settings;
su3_soa *auxbis_conf_acc;
su3_soa *glocal_staples;
tamat_soa *gipdot;

long ref_sizeh(void) { return sizeh; }
int ref_nd(int i) { int nd[4] = {nd0, nd1, nd2, nd3}; return nd[i]; }

/* compute_nnp_and_nnm_openacc (OpenAcc/geometry.c:14-100) */
void ref_nn(int *nnp, int *nnm)
{
    compute_nnp_and_nnm_openacc();
    memcpy(nnp, nnp_openacc, sizeof(nnp_openacc));
    memcpy(nnm, nnm_openacc, sizeof(nnm_openacc));
}

int ref_snum(int d0, int d1, int d2, int d3) { return snum_acc(d0, d1, d2, d3); }

/* calc_u1_phases (OpenAcc/backfield.c:180-187) */
void ref_phases(double *out, const int *map, const double *bf, double im_chem_pot, double charge)
{
    int nd[4] = {nd0, nd1, nd2, nd3};
    geom_par.xmap = map[0]; geom_par.ymap = map[1]; geom_par.zmap = map[2]; geom_par.tmap = map[3];
    geom_par.gnx = nd[map[0]]; geom_par.gny = nd[map[1]]; geom_par.gnz = nd[map[2]]; geom_par.gnt = nd[map[3]];
    bf_param b;
    b.ex = bf[0]; b.ey = bf[1]; b.ez = bf[2]; b.bx = bf[3]; b.by = bf[4]; b.bz = bf[5];
    calc_u1_phases((double_soa *)out, b, im_chem_pot, charge);
}

/* reductions (OpenAcc/fermionic_utilities.c:126-175) */
double ref_l2norm2(const void *v) { return l2norm2_global((const vec3_soa *)v); }
double ref_real_scal_prod(const void *a, const void *b) { return real_scal_prod_global((const vec3_soa *)a, (const vec3_soa *)b); }

/* rational approximations (RationalApprox/rationalapprox.c:83-236) */
int ref_ratapprox_read(RationalApprox *ra, char *name) { return rationalapprox_read_custom_nomefile(ra, name); }
double ref_ratapprox_evaluate(RationalApprox *ra, double x) { return rational_approx_evaluate(ra, x); }

/* ---- stout smearing and its force chain (OpenAcc/stouting.c, plaquettes.c,
 * su3_utilities.c) with the reference's own kernels */
#include "OpenAcc/stouting.h"
#include "OpenAcc/su3_utilities.h"
#include "OpenAcc/plaquettes.h"

void ref_stout_isotropic(const void *u, void *uout, double rho)
{
    compute_nnp_and_nnm_openacc();
    gl_stout_rho = rho;
    su3_soa *stap = malloc(8 * sizeof(su3_soa)), *aux = malloc(8 * sizeof(su3_soa));
    tamat_soa *ta = malloc(8 * sizeof(tamat_soa));
    stout_isotropic((const su3_soa *)u, (su3_soa *)uout, stap, aux, ta, 0);
    free(stap); free(aux); free(ta);
}

/* staples + rho*TA(U*staples) (plaquettes.c:196-259, su3_utilities.c:210-239) */
void ref_staples_and_qa(const void *u, void *stap, void *qa, double rho)
{
    compute_nnp_and_nnm_openacc();
    gl_stout_rho = rho;
    set_su3_soa_to_zero((su3_soa *)stap);
    calc_loc_staples_nnptrick_all_onlyferms((const su3_soa *)u, (su3_soa *)stap);
    RHO_times_conf_times_staples_ta_part((const su3_soa *)u, (const su3_soa *)stap, (tamat_soa *)qa, 0);
}

/* the body of compute_sigma_from_sigma_prime_backinto_sigma_prime
 * (fermion_force.c:95-160: staples, QA, compute_lambda, compute_sigma) */
void ref_sigma_prime_to_sigma(void *Sigma, void *Lambda, void *QA, const void *U, double rho)
{
    compute_nnp_and_nnm_openacc();
    gl_stout_rho = rho;
    su3_soa *tmp = malloc(8 * sizeof(su3_soa));
    set_su3_soa_to_zero(tmp);
    calc_loc_staples_nnptrick_all_onlyferms((const su3_soa *)U, tmp);
    RHO_times_conf_times_staples_ta_part((const su3_soa *)U, tmp, (tamat_soa *)QA, 0);
    compute_lambda((thmat_soa *)Lambda, (const su3_soa *)Sigma, (const su3_soa *)U, (const tamat_soa *)QA, tmp);
    compute_sigma((const thmat_soa *)Lambda, (const su3_soa *)U, (su3_soa *)Sigma, (const tamat_soa *)QA, tmp, 0);
    free(tmp);
}

void ref_unitarize(void *u) { unitarize_conf((su3_soa *)u); }

/* the tree-level Symanzik gauge force of calc_ipdot_gauge_soloopenacc_tlsm
 * (OpenAcc/ipdot_gauge.c:157-181) composed from the same routines in the
 * same order (ipdot_gauge.c itself pulls in the debug printers of DbgTools,
 * whose MPI buffers do not build here); single rank, no PAR_TEMP */
void ref_gauge_force_tlsm(const void *u, void *stap, void *ipdot)
{
    compute_nnp_and_nnm_openacc();
    set_su3_soa_to_zero((su3_soa *)stap);
    calc_loc_staples_nnptrick_all((const su3_soa *)u, (su3_soa *)stap);
    calc_loc_improved_staples_typeA_nnptrick_all((const su3_soa *)u, (su3_soa *)stap);
    calc_loc_improved_staples_typeB_nnptrick_all((const su3_soa *)u, (su3_soa *)stap);
    calc_loc_improved_staples_typeC_nnptrick_all((const su3_soa *)u, (su3_soa *)stap);
    conf_times_staples_ta_part((const su3_soa *)u, (const su3_soa *)stap, (tamat_soa *)ipdot);
}

/* calc_plaquette_soloopenacc (su3_measurements.c:191-211) and
 * calc_rettangolo_soloopenacc (rectangles.c:1890-1909) with their scratch */
double ref_plaquette(const void *u)
{
    compute_nnp_and_nnm_openacc();
    su3_soa *lp = malloc(2 * sizeof(su3_soa));
    dcomplex_soa *tr = malloc(2 * sizeof(dcomplex_soa));
    double v = calc_plaquette_soloopenacc((su3_soa *)u, lp, tr);
    free(lp); free(tr);
    return v;
}

double ref_rectangles(const void *u)
{
    compute_nnp_and_nnm_openacc();
    su3_soa *lp = malloc(2 * sizeof(su3_soa));
    dcomplex_soa *tr = malloc(2 * sizeof(dcomplex_soa));
    double v = calc_rettangolo_soloopenacc((const su3_soa *)u, lp, tr);
    free(lp); free(tr);
    return v;
}


/* mom_sum_mult and mom_exp_times_conf_soloopenacc (su3_utilities.c:241-307)
 * with a one-entry factor array, calc_momenta_action (su3_measurements.c:138-168) */
void ref_mom_sum_mult(void *mom, const void *ipdot, double fact)
{
    double f[1] = {fact};
    mom_sum_mult((thmat_soa *)mom, (const tamat_soa *)ipdot, f, 0);
}

void ref_mom_exp(void *u, const void *mom, double delta)
{
    double f[1] = {delta};
    mom_exp_times_conf_soloopenacc((su3_soa *)u, (const thmat_soa *)mom, f, 0);
}

double ref_momenta_action(const void *mom, int mu)
{
    double_soa *tr = malloc(sizeof(double_soa));
    double v = calc_momenta_action((const thmat_soa *)mom, tr, mu);
    free(tr);
    return v;
}

/* deltas_Omelyan of the compiled initialize_md_global_variables (md_parameters.c:11-52) */
void ref_md_deltas(double beta, int no_md, int gauge_scale, double t, double *out)
{
    md_param p;
    memset(&p, 0, sizeof p);
    p.no_md = no_md;
    p.gauge_scale = gauge_scale;
    p.t = t;
    act_params.beta = beta;
    initialize_md_global_variables(p);
    md_parameters = p;
    memcpy(out, deltas_Omelyan, sizeof deltas_Omelyan);
}

/* one sub-step of multistep_2MN_gauge_bloc (md_integrator.c:256-345) */
static void ref_gauge_bloc(void *u, void *stap, void *ipdot, void *mom, int idx)
{
    ref_gauge_force_tlsm(u, stap, ipdot);
    mom_sum_mult((thmat_soa *)mom, (const tamat_soa *)ipdot, deltas_Omelyan, idx);
    mom_exp_times_conf_soloopenacc((su3_soa *)u, (const thmat_soa *)mom, deltas_Omelyan, 4);
}

/* multistep_2MN_gauge (md_integrator.c:347-376) composed from the compiled
 * force, momentum and link routines, deltas from ref_md_deltas (md_integrator.c
 * itself needs the generated FP32 headers and does not build here) */
void ref_md_gauge(void *u, void *stap, void *ipdot, void *mom, double beta, int no_md, int gauge_scale, double t)
{
    double d[7];
    ref_md_deltas(beta, no_md, gauge_scale, t, d);
    ref_gauge_bloc(u, stap, ipdot, mom, 3);
    for (int md = 1; md < gauge_scale; md++) {
        ref_gauge_bloc(u, stap, ipdot, mom, 5);
        ref_gauge_bloc(u, stap, ipdot, mom, 6);
    }
    ref_gauge_bloc(u, stap, ipdot, mom, 5);
    ref_gauge_force_tlsm(u, stap, ipdot);
    mom_sum_mult((thmat_soa *)mom, (const tamat_soa *)ipdot, deltas_Omelyan, 3);
}

/* dM/dmu (kind 0) and d2M/dmu2 (kind 1) of machine direction dir, output even
 * (eo = 0) or odd (eo = 1) sites (Meas/baryon_number_utilities.c) */
void ref_dmu(int kind, int eo, int dir, const void *u, void *out, const void *in, const void *ph)
{
    compute_nnp_and_nnm_openacc();
    void (*f)(const su3_soa *, vec3_soa *, const vec3_soa *, const double_soa *) =
        kind == 0 ? (eo == 0 ? dM_dmu_eo[dir] : dM_dmu_oe[dir]) : (eo == 0 ? d2M_dmu2_eo[dir] : d2M_dmu2_oe[dir]);
    f((const su3_soa *)u, (vec3_soa *)out, (const vec3_soa *)in, (const double_soa *)ph);
}

/* idphase_dbz (magnetic_susceptibility_utilities.c:13-35) composed from the
 * compiled backfield.c routines it calls (that file needs generated headers) */
void ref_idphase_dbz(void *re, void *im, const int *map, double charge)
{
    int nd[4] = {nd0, nd1, nd2, nd3};
    geom_par.xmap = map[0]; geom_par.ymap = map[1]; geom_par.zmap = map[2]; geom_par.tmap = map[3];
    geom_par.gnx = nd[map[0]]; geom_par.gny = nd[map[1]]; geom_par.gnz = nd[map[2]]; geom_par.gnt = nd[map[3]];
    bf_param zero = {0, 0, 0, 0, 0, 0};
    bf_param bz1 = zero;
    bz1.bz += 1;
    calc_u1_phases_unb_no2pi((double_soa *)im, bz1, 0, charge);
    calc_u1_phases_unb_no2pi((double_soa *)re, zero, 0, charge);
    phase_diff_in_place((double_soa *)im, (double_soa *)re);
    mult_u1_phases((double_soa *)im, 2 * 3.14159265358979323846);
    set_double_soa_to_zero((double_soa *)re);
}
