/*
 * openstaple_amd.h -- C ABI of the MI355X-native OpenStaPLE fermion hot path.
 *
 * Drop-in boundary: every function below that carries a reference name keeps
 * the reference's signature (with `__restrict` dropped) and semantics, and
 * cites the declaration it replaces.  Field arguments are pointers to the
 * reference's structure-of-arrays types (OpenAcc/struct_c_def.h:16-58) laid out
 * exactly as the reference lays them out, with the lattice extents supplied at
 * run time through ostp_init() instead of compile-time macros.
 *
 * Pointer arguments may be either
 *   - host addresses registered with ostp_acc_create()/ostp_acc_copyin()
 *     (the replacement of `#pragma acc enter data create/copyin`); the library
 *     works on the device mirror, like the reference's `present(...)` kernels,
 *     and host copies change only through ostp_acc_update_self(); or
 *   - device addresses (ostp_malloc(), hipMalloc(), a torch tensor's data_ptr).
 * Any other address is an error (the reference's OpenACC runtime aborts on a
 * missing `present` datum; see ostp_set_abort_on_error()).
 *
 * All entry points are synchronous at return, like the reference's
 * non-async `acc kernels` regions.  Not re-entrant (the reference is not
 * either: it uses globals).
 */
#ifndef OPENSTAPLE_AMD_H
#define OPENSTAPLE_AMD_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference types (complete definitions: openstaple_amd_types.h) ----- */
typedef struct vec3_soa_t vec3_soa;         /* struct_c_def.h:16-20 */
typedef struct su3_soa_t su3_soa;           /* struct_c_def.h:35-42 */
typedef struct double_soa_t double_soa;     /* struct_c_def.h:25-27 */
typedef struct tamat_soa_t tamat_soa;       /* struct_c_def.h:45-51 */
typedef struct thmat_soa_t thmat_soa;       /* struct_c_def.h:52-58 */
typedef struct dcomplex_soa_t dcomplex_soa; /* struct_c_def.h:21-23 */
typedef struct vec3_soa_f_t vec3_soa_f;     /* generated sp_struct_c_def.h twin of vec3_soa */
typedef struct su3_soa_f_t su3_soa_f;       /* generated twin of su3_soa */
typedef struct float_soa_t float_soa;       /* generated twin of double_soa */
typedef struct ferm_param_t ferm_param;     /* Include/fermion_parameters.h:9-41 */
typedef struct RationalApprox_t RationalApprox; /* RationalApprox/rationalapprox.h:15-26 */
typedef struct inverter_package_t inverter_package; /* OpenAcc/inverter_package.h:12-29 */
typedef struct bf_param_t bf_param;         /* OpenAcc/backfield_parameters.h:5-9 */
typedef struct inv_tricks_t inv_tricks;     /* Include/inverter_tricks.h:4-11 */
typedef struct rep_info_t rep_info;         /* Include/rep_info.h:4-11 */
typedef struct defect_info_t defect_info;   /* OpenAcc/HPT_utilities.h:14-32 */
typedef struct md_param_t md_param;         /* OpenAcc/md_parameters.h:6-19 */
typedef struct action_param_t action_param; /* OpenAcc/action.h:6-18 */
typedef struct alloc_settings_t alloc_settings; /* OpenAcc/alloc_settings.h:4-20 */
typedef struct global_su3_soa_t global_su3_soa; /* struct_c_def.h:68-72: r0, r1, r2, no K plane */

/* Reference globals the replaced code reads or writes.  The library defines
 * them with the reference defaults; a host program that defines them itself
 * (the reference's setting_file_parser.c / inverter_wrappers.c) takes
 * precedence by ordinary symbol interposition. */
extern int verbosity_lv;                    /* Include/setting_file_parser.c */
extern int multishift_invert_iterations;    /* inverter_wrappers.h:14 */
extern double deltas_Omelyan[7];            /* md_parameters.h:21 */
extern float deltas_Omelyan_f[7];           /* md_parameters.h:22 */
extern md_param md_parameters;              /* md_parameters.h:28 */
extern action_param act_params;             /* action.h:22 (beta) */
extern int nMdInversionPerformed;           /* md_parameters.h:26 */
/* fields the RHMC update reads (OpenAcc/alloc_vars.h:20-76, alloc_settings.h:23,
 * Include/fermion_parameters.h:46); null until the host points them at its own */
extern su3_soa *conf_acc_bkp, *aux_conf_acc, *auxbis_conf_acc, *gstout_conf_acc_arr, *gconf_as_fermionmatrix;
extern thmat_soa *momenta;
extern tamat_soa *ipdot_acc;
extern vec3_soa *ferm_chi_acc, *ferm_phi_acc, *ferm_shiftmulti_acc, *kloc_r, *kloc_h, *kloc_s, *kloc_p, *k_p_shiftferm;
extern dcomplex_soa *local_sums;
extern double_soa *d_local_sums;
extern ferm_param *fermions_parameters;
extern alloc_settings alloc_info;

#define INVERTER_SUCCESS 1                  /* inverter_full.h:12-13 */
#define INVERTER_FAILURE 0
#define CONVERGENCE_CRITICAL 1              /* inverter_wrappers.h:10-11 */
#define CONVERGENCE_NONCRITICAL 0

/* ======================================================== runtime (new) === */
/* Lattice extents nd0..nd3 (d0 fastest, nd0 even) and whether su3_soa holds
 * the PAR_TEMP K plane.  Replaces the compile-time LOC_N0..3/NREPLICAS
 * (Mpi/geometry_multidev.h:17-27).  Returns 0 on success. */
int  ostp_init(int nd0, int nd1, int nd2, int nd3, int par_temp);
void ostp_finalize(void);
/* physical->logical direction map (geom_parameters xmap..tmap, geometry.h:33-37) */
int  ostp_set_direction_map(int xmap, int ymap, int zmap, int tmap);
/* sizeof checks for a host compiled against the reference headers */
int  ostp_check_sizes(size_t sizeof_vec3_soa, size_t sizeof_su3_soa, size_t sizeof_double_soa);
long ostp_sizeh(void);

/* OpenACC data-directive replacements (acc_create/acc_copyin/acc_update_*) */
int   ostp_acc_create(void *host, size_t bytes);
int   ostp_acc_copyin(void *host, size_t bytes);
int   ostp_acc_update_device(const void *host, size_t bytes);
int   ostp_acc_update_self(void *host, size_t bytes);
int   ostp_acc_delete(void *host);
void *ostp_acc_deviceptr(const void *host);
int   ostp_acc_is_present(const void *host, size_t bytes);

/* plain device memory */
void *ostp_malloc(size_t bytes);
void  ostp_free(void *dev);
int   ostp_memcpy_htod(void *dev, const void *host, size_t bytes);
int   ostp_memcpy_dtoh(void *host, const void *dev, size_t bytes);
int   ostp_memcpy_dtod(void *dst, const void *src, size_t bytes);
int   ostp_memset(void *dev, int value, size_t bytes);
int   ostp_synchronize(void);
/* Entry points return synchronised by default (the reference's non-async
 * `acc kernels`).  With 0, Dslash, M^dag M and the vector BLAS entry points
 * only queue their kernels on the library stream (as the reference's
 * async(n) variants do) and ostp_synchronize() waits; solvers, reductions
 * and everything that returns a value still synchronise.  Returns the
 * previous setting. */
int   ostp_set_sync_at_return(int on);

/* errors: by default a bad pointer or a HIP failure prints a message and aborts
 * (as the OpenACC runtime does for a missing `present` datum); with 0 the call
 * returns (INVERTER_FAILURE for solvers) and the message is kept here. */
void        ostp_set_abort_on_error(int abort_on_error);
const char *ostp_last_error(void);
void        ostp_clear_error(void);

/* inverter_tricks (Include/inverter_tricks.h:4-11) as seen by the solvers */
void ostp_set_inverter_tricks(int singlePInvAccelMultiInv, int useMixedPrecision,
                              double mixedPrecisionDelta, int restartingEvery);

/* ======================================================== geometry ======= */
/* compute_nnp_and_nnm_openacc (OpenAcc/geometry.c:14-100): fills host arrays
 * laid out [sizeh][4][2] like nnp_openacc/nnm_openacc (geometry.h:59-60). */
int ostp_compute_nnp_and_nnm(int *nnp, int *nnm);

/* calc_u1_phases (OpenAcc/backfield.c:180-187): staggered phases, e.m. field,
 * imaginary chemical potential and antiperiodic time, rebounded, times 2 pi.
 * Computed on the device, bit-identical to the reference's double arithmetic. */
void calc_u1_phases(double_soa *phases, bf_param bf_pars, double im_chem_pot, double ferm_charge);
/* convert_double_to_float_real_soa on all 8 phase arrays (float_double_conv.h:31-34) */
void ostp_phases_to_float(const double_soa *phases, float_soa *phases_f);

/* ======================================================== Dslash ========= */
/* fermion_matrix.h:20-28 */
void acc_Deo(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void acc_Doe(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
/* fermion_matrix.h:95-106 */
void fermion_matrix_multiplication(const su3_soa *u, vec3_soa *out, const vec3_soa *in,
                                   vec3_soa *temp1, ferm_param *pars);
void fermion_matrix_multiplication_shifted(const su3_soa *u, vec3_soa *out, const vec3_soa *in,
                                           vec3_soa *temp1, ferm_param *pars, double shift);
/* generated FP32 twins (sp_fermion_matrix.h: names + "_f", float data) */
void acc_Deo_f(const su3_soa_f *u, vec3_soa_f *out, const vec3_soa_f *in, const float_soa *backfield);
void acc_Doe_f(const su3_soa_f *u, vec3_soa_f *out, const vec3_soa_f *in, const float_soa *backfield);
void fermion_matrix_multiplication_shifted_f(const su3_soa_f *u, vec3_soa_f *out, const vec3_soa_f *in,
                                             vec3_soa_f *temp1, ferm_param *pars, double shift);

/* ======================================================== BLAS =========== */
/* fermionic_utilities.h:45-118 (single rank) */
double l2norm2_global(const vec3_soa *in_vect1);
double real_scal_prod_global(const vec3_soa *in_vect1, const vec3_soa *in_vect2);
/* fermionic_utilities.c:126-145: sum over half-sites of conj(in1) . in2,
 * returned as the reference's d_complex (C99 double complex) */
double _Complex scal_prod_global(const vec3_soa *in_vect1, const vec3_soa *in_vect2);
/* the same, split into re/im for callers without a complex type (ctypes) */
void   ostp_scal_prod_global(const vec3_soa *in_vect1, const vec3_soa *in_vect2, double *re, double *im);
void combine_in1xfactor_plus_in2(const vec3_soa *in_vect1, const double factor,
                                 const vec3_soa *in_vect2, vec3_soa *out);
void combine_add_factor_x_in2_to_in1(vec3_soa *in1, const vec3_soa *in2, double factor);
void combine_in1xferm_mass_minus_in2(const vec3_soa *in_vect1, double ferm_mass2, vec3_soa *in_vect2);
void combine_in1_minus_in2(const vec3_soa *in_vect1, const vec3_soa *in_vect2, vec3_soa *out);
void assign_in_to_out(const vec3_soa *in_vect1, vec3_soa *out);
void set_vec3_soa_to_zero(vec3_soa *fermion);
/* float_double_conv.h:14-45 */
void convert_double_to_float_vec3_soa(const vec3_soa *d_var, vec3_soa_f *f_var);
void convert_float_to_double_vec3_soa(const vec3_soa_f *f_var, vec3_soa *d_var);
/* converts all 8 su3_soa of a gauge field (the reference call sites pass conf[0..7]) */
void ostp_convert_double_to_float_gauge(const su3_soa *d_var, su3_soa_f *f_var);

/* ======================================================== solvers ======== */
/* inverter_full.h:20-31 */
int ker_invert_openacc(const su3_soa *u, ferm_param *pars, vec3_soa *solution, const vec3_soa *in,
                       double res, vec3_soa *loc_r, vec3_soa *loc_h, vec3_soa *loc_s, vec3_soa *loc_p,
                       const int max_cg, double shift, int *cg_return);
/* No counterpart in the reference.  ker_invert_openacc for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): the same arguments, kernels and results
 * (bit-identical), with the reduction partials and the solver state in a
 * static device array of the library instead of the buffers ostp_init
 * allocated; the call itself allocates nothing.  Limit: lattices of at most
 * 2^23 half-sites (64^4); larger ones are rejected with INVERTER_FAILURE and
 * a recorded error.  The array is one per GPU and keeps the last call's
 * contents; calls that overlap on one GPU race (the library is not
 * re-entrant). */
int ker_invert_openacc_noalloc(const su3_soa *u, ferm_param *pars, vec3_soa *solution, const vec3_soa *in,
                               double res, vec3_soa *loc_r, vec3_soa *loc_h, vec3_soa *loc_s,
                               vec3_soa *loc_p, const int max_cg, double shift, int *cg_return);
/* inverter_mixedp.h:15-22 */
int inverter_mixed_precision(inverter_package ip, ferm_param *pars, vec3_soa *solution,
                             const vec3_soa *in, double res, const int max_cg, double shift,
                             int *cg_return);
/* inverter_multishift_full.h:15-27 */
int multishift_invert(const su3_soa *u, ferm_param *pars, RationalApprox *approx, vec3_soa *out,
                      const vec3_soa *in, double residuo, vec3_soa *loc_r, vec3_soa *loc_h,
                      vec3_soa *loc_s, vec3_soa *loc_p, vec3_soa *shiftferm, const int max_cg,
                      int *cg_return);
/* inverter_multishift_full.h:30-33 */
void recombine_shifted_vec3_to_vec3(const vec3_soa *in_shifted, const vec3_soa *in, vec3_soa *out,
                                    const RationalApprox *approx);
/* inverter_wrappers.h:16-34 */
void convergence_messages(int conv_importance, int inverter_status);
int inverter_wrapper(inverter_package ip, ferm_param *pars, vec3_soa *out, const vec3_soa *in,
                     double res, int max_cg, double shift, int convergence_importance);
int inverter_multishift_wrapper(inverter_package ip, ferm_param *pars, RationalApprox *approx,
                                vec3_soa *out, const vec3_soa *in, double res, int max_cg,
                                int convergence_importance);

/* Same solvers with the diagnostics the reference only prints: iteration
 * count, FP64 true residual ||b - M x|| / ||b||, reliable-update count, per-shift
 * residuals; explicit restart / reliable-update parameters instead of the
 * inverter_tricks global; `single` runs the CG-M in FP32 (multishift_invert_f). */
int ostp_cg_solve(const su3_soa *u, ferm_param *pars, vec3_soa *solution, const vec3_soa *in, double res,
                  vec3_soa *loc_r, vec3_soa *loc_h, vec3_soa *loc_s, vec3_soa *loc_p, int max_cg,
                  double shift, int restarting_every, int *cg_return, double *true_res);
int ostp_mixed_solve(inverter_package ip, ferm_param *pars, vec3_soa *solution, const vec3_soa *in,
                     double res, int max_cg, double shift, double mixed_delta, int *cg_return,
                     int *reliable_updates, double *true_res);
/* No reference counterpart: the CG-M keeps the base residuals of a window of
 * iterations and applies every shift's x_k / p_k update once per window
 * (the scalar recurrence is unchanged).  w = 0: window from the free memory
 * (up to 64; per-iteration updates below 8), 1: per-iteration updates,
 * 2..64: that window.  Returns the previous setting. */
int ostp_set_multishift_window(int w);
int ostp_multishift_solve(const su3_soa *u, ferm_param *pars, RationalApprox *approx, vec3_soa *out,
                          const vec3_soa *in, double residuo, vec3_soa *loc_r, vec3_soa *loc_h,
                          vec3_soa *loc_s, vec3_soa *loc_p, vec3_soa *shiftferm, int max_cg, int single,
                          int *cg_return, double *shift_res);

/* ======================================================== stout ========== */
/* rho of the fermion / topological stouting (OpenAcc/action.h:23) */
extern double gl_stout_rho;
extern double gl_topo_rho;
/* stout parameters: act_params.stout_rho/stout_steps/topo_* of the reference
 * input (Include/setting_file_parser.c:416) and the staple normalisation
 * C_ZERO of the gauge action (tlsm != 0: 5/3, tree-level Symanzik; else 1) */
void ostp_set_stout(double rho, int steps, double topo_rho, int topo_steps, int topo_action, int tlsm);
/* stouting.h:13-44 */
void stout_wrapper(const su3_soa *tconf_acc, su3_soa *tstout_conf_acc_arr, const int istopo);
void stout_isotropic(const su3_soa *u, su3_soa *uprime, su3_soa *local_staples, su3_soa *auxiliary,
                     tamat_soa *tipdot, const int istopo);
void exp_minus_QA_times_conf(const su3_soa *tu, const tamat_soa *QA, su3_soa *tu_out, su3_soa *exp_aux);
void compute_lambda(thmat_soa *L, const su3_soa *SP, const su3_soa *U, const tamat_soa *QA, su3_soa *TMP);
void compute_sigma(const thmat_soa *L, const su3_soa *U, su3_soa *S, const tamat_soa *QA, su3_soa *TMP,
                   const int istopo);
/* fermion_force.h:15-20 */
void compute_sigma_from_sigma_prime_backinto_sigma_prime(su3_soa *Sigma, thmat_soa *Lambda, tamat_soa *QA,
                                                          const su3_soa *U, su3_soa *TMP, const int istopo);
/* plaquettes.h / su3_utilities.h */
void calc_loc_staples_nnptrick_all_onlyferms(const su3_soa *u, su3_soa *loc_stap);
void RHO_times_conf_times_staples_ta_part(const su3_soa *u, const su3_soa *loc_stap, tamat_soa *tipdot,
                                          int istopo);
/* su3_utilities.h:25 */
void unitarize_conf(su3_soa *u);
int  ostp_stout_steps(void);

/* ======================================================== fermion force == */
/* fermion_force_utilities.h:16-42 */
void direct_product_of_fermions_into_auxmat(const vec3_soa *loc_s, const vec3_soa *loc_h, su3_soa *aux_u,
                                            const RationalApprox *approx, int iter);
void multiply_backfield_times_force(ferm_param *tpars, const su3_soa *auxmat, su3_soa *pseudo_ipdot);
void multiply_conf_times_force_and_take_ta_nophase(const su3_soa *u, const su3_soa *auxmat, tamat_soa *ipdot);
void ker_openacc_compute_fermion_force(const su3_soa *u, su3_soa *aux_u, const vec3_soa *in_shiftmulti,
                                       vec3_soa *loc_s, vec3_soa *loc_h, ferm_param *tpars);
/* fermion_force.h:24-36 (STOUT_FERMIONS build; stout steps from ostp_set_stout) */
void fermion_force_soloopenacc(su3_soa *tconf_acc, su3_soa *tstout_conf_acc_arr, su3_soa *gl3_aux,
                               tamat_soa *tipdot_acc, ferm_param *tfermion_parameters, int tNDiffFlavs,
                               const vec3_soa *ferm_in_acc, double res, su3_soa *taux_conf_acc,
                               vec3_soa *tferm_shiftmulti_acc, inverter_package ipt, const int max_cg);

/* ============================================ even-odd full inversion ==== */
/* fermionic_utilities.c:379-401: in2 = fact1*in1 - in2 */
void combine_in1_x_fact1_minus_in2_back_into_in2(const vec3_soa *in1, double fact1, vec3_soa *in2);
/* fermionic_utilities.c:403-419: out = fact*(in1 - in2) */
void combine_in1_minus_in2_allxfact(const vec3_soa *in1, const vec3_soa *in2, double fact, vec3_soa *out);
/* Meas/ferm_meas.c:50-71: (out_e, out_o) = M^-1 (in_e, in_o) for the full
 * staggered matrix M = [[m, Deo], [Doe, m]] through the even-odd system
 * (m^2 - Deo Doe) out_e = m in_e - Deo in_o (inverter_wrapper, critical),
 * out_o = (in_o - Doe out_e) / m.  phi_e, phi_o are scratch. */
void eo_inversion(inverter_package ip, ferm_param *tfermions_parameters, double res, int max_cg,
                  vec3_soa *in_e, vec3_soa *in_o, vec3_soa *out_e, vec3_soa *out_o, vec3_soa *phi_e,
                  vec3_soa *phi_o);

/* ================================== parallel tempering (PAR_TEMP builds) === */
/* OpenAcc/HPT_utilities.c:30-310.  K = 1 on every link, K = c_r on the links
 * of direction def_axis (mapped through ostp_set_direction_map) whose site
 * lies on the defect: coordinate nd-1 along def_axis, [0, def_vec[i]) along
 * the three orthogonal directions (x-y-z-t order).  When defect_info_config
 * is 0 the Delta-S swap ranges in *def are filled as well (single rank).
 * Needs ostp_init(..., par_temp = 1): K is the plane after row 2 of each
 * su3_soa. */
void init_k(su3_soa *conf, double c_r, int def_axis, int *def_vec, defect_info *def, int defect_info_config);

/* Replica label bookkeeping of manage_replica_swaps (HPT_utilities.c:722-1003)
 * for replicas held by one process.  Labels: replica index r carries label
 * label[r]; replica r runs with coefficient cr_vec[label[r]].
 *   action(cr_of_replica, label, S_of_label, user) must store, for every
 *     replica r, the defect action of replica r evaluated with coefficient
 *     cr_of_replica[r] at S_of_label[label[r]] (compute_S_of_replicas);
 *   uniform(user) returns a uniform number in [0,1) (the reference's casuale()).
 * Random numbers are drawn in the reference's order: one for the swap order,
 * then one per proposed pair whose Delta S is not negative.  On return every
 * replica r runs with cr_vec[label[r]] again (accepted pairs exchanged their
 * labels); swap_num, all_swap_vet[lab] and acceptance_vet[lab] are
 * incremented as in the reference (indexed by the lower label of a pair). */
typedef void (*ostp_pt_action_fn)(const double *cr_of_replica, const int *label, double *S_of_label, void *user);
typedef double (*ostp_pt_uniform_fn)(void *user);
void ostp_pt_manage_replica_swaps(rep_info *hpt_params, ostp_pt_action_fn action, ostp_pt_uniform_fn uniform,
                                  void *user, int *swap_num, int *all_swap_vet, int *acceptance_vet);
/* main.c:341-342: label[r] = r */
void ostp_pt_init_labels(rep_info *hpt_params);
/* rep_info.c:8-14 (get_index_of_pbc_replica): replica index carrying label 0, -1 if none */
int ostp_pt_index_of_label0(const rep_info *hpt_params);
/* metro_SWAP_worldmaster (HPT_utilities.c:704-719) with the uniform draw
 * passed in: accept if Delta_S < 0, else if u < exp(-Delta_S). */
int ostp_pt_metropolis(double delta_s_swap, double u);
/* Partner label of `lab` in an even/odd stage (HPT_utilities.c:769); -1 if
 * the label sits out this stage. */
int ostp_pt_twin_label(int lab, int swap_order, int nreplicas);

/* ============================================= batched solves (new) ===== */
/* No reference counterpart: the reference solves one right-hand side per
 * call.  One system of a batch: the even-odd system (m^2 - Deo Doe) x = b on
 * its own lattice (extents nd, K plane iff par_temp), solved by the
 * inverter_mixed_precision algorithm (OpenAcc/inverter_mixedp.c:41-181:
 * FP32 CG, FP64 reliable updates, SAFETY 0.9, final FP64 true residual).
 * Pointers as everywhere else (registered host or device addresses). */
typedef struct ostp_batch_system_t {
    int nd[4];
    int par_temp;
    const su3_soa *u;          /* FP64 links: reliable updates, final residual */
    const su3_soa_f *u_f;      /* FP32 links: inner iterations */
    const double_soa *phases;
    const float_soa *phases_f;
    const vec3_soa *in;        /* b */
    vec3_soa *out;             /* x: initial guess on entry, solution on return */
    double mass;
} ostp_batch_system;
/* Solve all systems together: every kernel is batched over systems, each
 * system keeps its scalars on the device, and iters_per_graph iterations are
 * captured once into a HIP graph that is replayed until all systems are
 * done (iters_per_graph = 0: the same kernels launched directly, for
 * profilers and debuggers that do not follow stream capture).  The Dslash
 * tile is autotuned per batch signature on first use.
 * cg_return[k], ru_return[k] (reliable updates) and true_res[k] (FP64
 * ||b - M x|| / ||b||) may be NULL.  Returns the number of systems whose true
 * residual is <= res (the ostp_init lattice is not used by this call, but
 * ostp_init must have been called once). */
int ostp_batched_mixed_solve(int nsys, const ostp_batch_system *systems, double res, int max_cg,
                             double mixed_delta, int iters_per_graph, int *cg_return, int *ru_return,
                             double *true_res);
/* The same with FP16 storage in the inner solver (links rows 0,1 and phases
 * as (cos, sin), inner vectors; FP32 arithmetic, FP32 solution accumulator,
 * FP64 reliable updates).  The inner vectors are scaled so that their
 * entries stay O(1): at the start and at every reliable update the FP64
 * residual is scaled to mean |entry|^2 = 1 and the search direction by the
 * ratio of the new and old scales; delta carries the same factor, so the
 * recurrences are the reference's in exact arithmetic.  No reference
 * counterpart (the reference's inner precision is FP32). */
int ostp_batched_half_solve(int nsys, const ostp_batch_system *systems, double res, int max_cg,
                            double mixed_delta, int iters_per_graph, int *cg_return, int *ru_return,
                            double *true_res);
/* Rational solve with reduced-precision inner work: FP32 CG-M (FP32 Dslash and
 * vectors, the FP64 shift recurrences of multishift_invert) with FP64
 * reliable updates of the base residual (r = b - M x0 whenever |r|^2 has
 * fallen by 0.1 since the last update; FP32 solution accumulators added to the
 * FP64 solutions), then every shift refined, all together, by the batched
 * solver (half_inner: FP16 storage, else FP32; FP64 reliable updates) from
 * its CG-M solution.
 * Buffers: ip.u, ip.u_f, ip.ferm_shift_temp_f (approx_order FP32 vectors),
 * ip.loc_{r,h,s,p}_f, ip.out_f, ip.loc_{r,h,s,p} (FP64 work); pars->phases,
 * pars->phases_f; out holds approx_order vec3_soa.  cg_multishift: CG-M
 * iterations; cg_refine[k], true_res[k] per shift (may be NULL).  Returns the number of shifts whose FP64 true residual
 * ||b - (M + RA_b[k]) x_k|| / ||b|| is <= res.  No reference counterpart (the
 * reference's single-precision accelerated path is inverter_multishift_wrapper
 * with singlePInvAccelMultiInv). */
int ostp_multishift_mixed_solve(inverter_package ip, ferm_param *pars, RationalApprox *approx, vec3_soa *out,
                                const vec3_soa *in, double res, int max_cg, int half_inner, int *cg_multishift,
                                int *cg_refine, double *true_res);
/* The solve inverter_multishift_wrapper performs (heat bath, MD force and
 * acceptance of UPDATE_SOLOACC_UNOSTEP_VERSATILE, fermion_force_soloopenacc):
 * mode 0 (default) as the reference selects it with inverter_tricks; 1 the
 * reduced-precision rational solve of ostp_multishift_mixed_solve with FP32
 * refinement inner solves; 2 the same with FP16 inner solves.  With
 * force_only, modes 1 and 2 apply to the solves inside
 * fermion_force_soloopenacc (the MD force) and every other call keeps mode 0.
 * Modes 1 and 2 use the package's FP32 links and vectors and the flavour's
 * FP32 phases when they are set, library buffers converted from the FP64
 * ones when not.  The returned iteration count adds the refinement
 * iterations of every shift.  Returns the previous mode (-1 and a recorded
 * error for a bad one). */
int ostp_set_multishift_precision(int mode, int force_only);
/* Iterations per HIP graph of the batched solves the library runs itself
 * (the refinement of ostp_multishift_mixed_solve, ostp_chiral_condensate);
 * default 16, 0 launches the kernels directly (for tracers that do not
 * follow graph launches).  Returns the previous value, -1 and a recorded
 * error outside 0..4096. */
int ostp_set_batched_graph_iterations(int n);
/* No reference counterpart.  fermion_force_soloopenacc solves the rational
 * systems of all its (flavour, pseudofermion) sources before forming the
 * forces; when they are FP64 solves (no reduced-precision or single-precision
 * accelerated mode) on one operator up to the mass (the same U(1) phases:
 * one array, or equal values), up to four run as one CG-M each whose
 * iterations share the two hops (each link applied to every source's
 * spinor); UPDATE_SOLOACC_UNOSTEP_VERSATILE does the same for its heat-bath
 * and acceptance solves and for the flavours' power iterations (eigenvalue
 * bounds).  Each source's iterations, solutions and convergence are those of
 * its own multishift_invert (power method).  on = 0 runs them one by one;
 * returns the previous setting (default 1). */
int ostp_set_multishift_batching(int on);
/* Text of the autotuner's measurements so far (candidates and choice per
 * batch signature); returns its length, copies up to len-1 bytes into buf. */
int ostp_autotune_report(char *buf, int len);
/* Fix the batched Dslash tile instead of autotuning it: blk = 64, 128 or 256
 * threads, order bit 0 t-fastest tiles, bit 1 hardware workgroup order, bit 2
 * systems interleaved (the workgroups of one tile for every system follow
 * each other, so systems on one gauge field share its links through L2),
 * bit 3 chunked hops (one thread applies each link to up to eight systems on
 * one gauge field, in every precision); blk = 0
 * returns to the autotuner.  Bits 1 and 2 do not change results (reductions
 * are per tile in a fixed order).  Returns the previous setting
 * (blk * 16 + order, 0 for the autotuner), -1 and a recorded error for a bad
 * one. */
int ostp_set_batch_tile(int blk, int order);

/* ============ molecular dynamics (su3_utilities.h, su3_measurements.h, md_integrator.h) */
/* su3_utilities.h:53-57: mom[l] += fact * ipdot[l] in the sense
 * P -= i fact TA (c01, c02, c12) and P_kk += fact TA_kk, for all 8 link arrays,
 * fact = factor[id_factor] read on the host (the reference's deltas_Omelyan) */
void mom_sum_mult(thmat_soa *mom, const tamat_soa *ipdot, const double *factor, int id_factor);
/* su3_utilities.h:58-61: U <- exp(i factor[id_factor] P) U on rows 0 and 1,
 * re-unitarised (row 2 is not written, as in the reference) */
void mom_exp_times_conf_soloopenacc(su3_soa *tconf_acc, const thmat_soa *mom, const double *factor, int id_factor);
/* su3_measurements.h: sum over the sites of link array mu (0..7) of Tr(P^2)/2;
 * tr_local (work space of the reference) is not written */
double calc_momenta_action(const thmat_soa *mom, double_soa *tr_local, const int mu);
/* su3_measurements.h: P <- -P on all 8 link arrays */
void invert_momenta(thmat_soa *mom);
/* md_parameters.h:24: deltas_Omelyan[0..6] (and the float copies) from the
 * MD parameters and act_params.beta; md_parameters itself is set by the host */
void initialize_md_global_variables(md_param md_params);
/* md_integrator.h: the gauge sub-integrator (2MN with md_parameters.gauge_scale
 * steps over half an outer step).  Each force evaluation is fused with its
 * momentum step; local_staples receives the staples of the last evaluation. */
void multistep_2MN_gauge(su3_soa *tconf_acc, su3_soa *local_staples, tamat_soa *tipdot, thmat_soa *tmomenta);
/* md_integrator.h (STOUT_FERMIONS build): one MD trajectory of
 * md_parameters.no_md 2MN steps with the fermion force of every flavour and the
 * gauge sub-integrator.  Not supported (error): act_params.topo_action == 1,
 * md_parameters.extrapolateInvsForce. */
void multistep_2MN_SOLOOPENACC(tamat_soa *tipdot_acc, su3_soa *tconf_acc, su3_soa *tstout_conf_acc_arr,
                               su3_soa *tauxbis_conf_acc, su3_soa *taux_conf_acc, ferm_param *tfermions_parameters,
                               int tNDiffFlavs, vec3_soa *ferm_in_acc, vec3_soa *tferm_shiftmulti_acc,
                               inverter_package ip, thmat_soa *tmomenta, dcomplex_soa *local_sums, double res,
                               const int max_cg);

/* ========================= fermion measurements (Meas/ferm_meas.h and helpers) */
/* baryon_number_utilities.h:13-31: dM/dmu and d2M/dmu2 hop operators of one
 * machine direction (arrays indexed by the direction of time), output even
 * (eo) or odd (oe) sites; 0.5 (U in(x+mu) +/- U^+ in(x-mu)) with the phases */
void dM_dmu_eo0(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void dM_dmu_eo1(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void dM_dmu_eo2(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void dM_dmu_eo3(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void dM_dmu_oe0(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void dM_dmu_oe1(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void dM_dmu_oe2(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void dM_dmu_oe3(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void d2M_dmu2_eo0(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void d2M_dmu2_eo1(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void d2M_dmu2_eo2(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void d2M_dmu2_eo3(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void d2M_dmu2_oe0(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void d2M_dmu2_oe1(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void d2M_dmu2_oe2(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
void d2M_dmu2_oe3(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *backfield);
extern void (*dM_dmu_eo[4])(const su3_soa *, vec3_soa *, const vec3_soa *, const double_soa *);
extern void (*dM_dmu_oe[4])(const su3_soa *, vec3_soa *, const vec3_soa *, const double_soa *);
extern void (*d2M_dmu2_eo[4])(const su3_soa *, vec3_soa *, const vec3_soa *, const double_soa *);
extern void (*d2M_dmu2_oe[4])(const su3_soa *, vec3_soa *, const vec3_soa *, const double_soa *);
/* field_times_fermion_matrix.h: Deo / Doe with every hop's phase multiplied by
 * the complex field field_re + i field_im (e.g. i dphase/db_z) */
void acc_Deo_wf(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *phases,
                const double_soa *field_re, const double_soa *field_im);
void acc_Doe_wf(const su3_soa *u, vec3_soa *out, const vec3_soa *in, const double_soa *phases,
                const double_soa *field_re, const double_soa *field_im);
/* magnetic_susceptibility_utilities.h: i dphase/db_z of the fermion's phases
 * (re = 0; im from the unbounded phases at b_z = 1 and 0) */
void idphase_dbz(double_soa *idphi_dbz_re, double_soa *idphi_dbz_im, ferm_param *fpar);
/* ferm_meas.c:230-460, one flavour, all noise vectors' inversions batched:
 * out[12 k + 2 q + (0, 1)] = (re, im) x degeneracy / (4 V) for vector k and
 * q = 0 chiral condensate <r, M^-1 r>, 1 quark number <r, dM/dmu M^-1 r>,
 * 2 magnetisation -<r, dM/db M^-1 r> (with pars->mag_re/mag_im, else NaN),
 * 3 -<r, M^-2 r> (k < n_chiral_susc), 4 <r, d2M/dmu2 M^-1 r> and
 * 5 -<r, (dM/dmu M^-1)^2 r> (k < n_qn_susc); entries not measured are NaN.
 * r: Z2 noise from streams 2k (even) and 2k+1 (odd) of seed.  Returns nvec
 * less the number of inversions that did not converge (floored at 0; nvec
 * when every inversion converged), 0 on an error. */
int ostp_fermion_measures(const su3_soa *u, const su3_soa_f *u_f, ferm_param *pars, int nvec, int n_chiral_susc,
                          int n_qn_susc, uint64_t seed, double res, int max_cg, int half_inner, double *out,
                          int *cg_return);

/* ====================================== RHMC update (update_versatile.h) */
/* update_versatile.c:44-580 (STOUT_FERMIONS, double-precision MD): one RHMC
 * trajectory on tconf_acc with the reference's steps, actions and Metropolis
 * test; returns acc + 1 if accepted (or metro == 0), else acc.  Reads
 * act_params.beta, md_parameters, alloc_info (NDiffFlavs, maxApproxOrder),
 * fermions_parameters (masses, phases, mother approximations) and the field
 * globals above.  Not supported (error): md_parameters.singlePrecMD,
 * act_params.topo_action, the NORANDOM and reversibility debug modes. */
int UPDATE_SOLOACC_UNOSTEP_VERSATILE(su3_soa *tconf_acc, double res_metro, double res_md, int id_iter, int acc,
                                     int metro, int max_cg);

/* No reference counterpart: the random streams of the update.  Momenta,
 * pseudofermions and power-method starts draw stream first_stream, first_stream
 * + 1, ... of the counter-based generator with this seed; the Metropolis
 * uniform comes from fn(ctx) when set (e.g. the host's casuale), else from the
 * same seed. */
typedef double (*ostp_uniform_fn)(void *ctx);
void ostp_set_rng_seed(uint64_t seed, uint32_t first_stream);
void ostp_set_metropolis_uniform(ostp_uniform_fn fn, void *ctx);
/* the actions and the Metropolis decision of the last update (the values the
 * reference prints) */
typedef struct ostp_rhmc_info_t {
    int iteration;
    int accepted;
    int md_multishift_iterations;
    double action_in, action_mom_in, action_ferm_in;    /* gauge action beta/3 (c0 P + c1 R) */
    double action_fin, action_mom_fin, action_ferm_fin;
    double delta_S, p1, p2;                             /* p2 = -1: no draw (delta_S < 0) */
} ostp_rhmc_info;
int ostp_rhmc_last_info(ostp_rhmc_info *out);

/* ============================== field files (OpenAcc/io.h, DbgTools/dbgtools.h) */
/* Host arrays, as in the reference (a device address is staged through a host
 * copy); the lattice comes from ostp_init or, for programs without a GPU,
 * from ostp_set_lattice_host.  Formats, record order, byte order and return
 * codes are the reference's (0 ok, 1 file missing, 2 malformed). */
int ostp_set_lattice_host(int nd0, int nd1, int nd2, int nd3, int par_temp);
/* io.c:35-108, io.c:111-207 */
int print_su3_soa_ASCII(const global_su3_soa *conf_rw, const char *nomefile, int conf_id_iter,
                        const double_soa *back_phases);
int read_su3_soa_ASCII(global_su3_soa *conf, const char *nomefile, int *conf_id_iter);
/* io.c:257-412, io.c:415-547: LIME records ildg-format, MD_traj, input-file,
 * ildg-binary-data (big-endian FP64, t z y x order, directions x y z t) */
int print_su3_soa_ildg_binary(const global_su3_soa *conf, const char *nomefile, int conf_id_iter);
int read_su3_soa_ildg_binary(global_su3_soa *conf, const char *nomefile, int *conf_id_iter);
/* dbgtools.c:1193-1236 */
/* io.c:553-614 (single rank): visit the global sites in ILDG file order
 * (t, z, y, x, x fastest; the four directions x, y, z, t of each site, or
 * only the even sites and one call each with scalar_even_mode) and call
 * single_element_rw(half-site index in the machine layout, parity, machine
 * direction, datastruct, 1 if this host is little endian, fp).  Extents and
 * direction map as for the field files. */
void rw_iterate_on_global_sites_lx_xyzt_axis_ordering(
    void (*single_element_rw)(const int idxh_machine, const int parity, const int dirmachine,
                              const void *data, const int conf_machine_endianness_disagreement, FILE *fp),
    const void *datastruct, FILE *fp, const int scalar_even_mode);
void print_vec3_soa(const vec3_soa *fermion, const char *nomefile);
int  read_vec3_soa(vec3_soa *fermion, const char *nomefile);
/* save_conf_wrapper / read_conf_wrapper (io.h:51-150) on the library's
 * su3_soa[8] (device or registered address; the K plane is not written) */
int ostp_save_conf(const su3_soa *conf, const char *nomefile, int conf_id_iter, int use_ildg);
int ostp_read_conf(su3_soa *conf, const char *nomefile, int *conf_id_iter, int use_ildg);
extern char input_file_str[300 * 500];      /* Include/setting_file_parser.h:11 */

/* ====================== gauge force and action (ipdot_gauge.h, su3_measurements.h, rectangles.h) */
/* ipdot_gauge.h:15-17: local_staples = C_ZERO x plaquette staples (+ C_ONE x
 * rectangle staples of types A, B, C for the tree-level Symanzik action set by
 * ostp_set_stout's tlsm flag), times the defect coefficients with PAR_TEMP;
 * tipdot = TA(U * staple).  One fused kernel. */
void calc_ipdot_gauge_soloopenacc(const su3_soa *tconf_acc, su3_soa *local_staples, tamat_soa *tipdot);
/* su3_measurements.h:14 and rectangles.h: sum over sites and planes of
 * Re tr of the plaquettes / of the 2x1 and 1x2 rectangles (with the PAR_TEMP
 * defect coefficients).  The scratch arguments are not written. */
double calc_plaquette_soloopenacc(const su3_soa *tconf_acc, su3_soa *local_plaqs, dcomplex_soa *tr_local_plaqs);
double calc_rettangolo_soloopenacc(const su3_soa *tconf_acc, su3_soa *local_plaqs, dcomplex_soa *tr_local_plaqs);
/* No reference counterpart: device microseconds per launch of the gauge-force
 * kernel (staples when stap is not null, ipdot, and the fused momentum step
 * mom_sum_mult(mom, ipdot, fact) when mom is not null), averaged over reps */
double ostp_time_gauge_force(const su3_soa *u, su3_soa *stap, tamat_soa *ipdot, thmat_soa *mom, double fact,
                             int reps);
/* No reference counterpart: which kernels form the gauge force when the
 * staples are not asked for (the MD's sub-steps, ostp_time_gauge_force with
 * stap null): 1 (default) one launch per plane with the plane's links in LDS,
 * where every extent is a multiple of 4; 0 the per-link kernel.  Returns the
 * previous setting. */
int ostp_set_gauge_force_kernel(int planes);
/* No reference counterpart: the kernel that accumulates the fermion force's
 * outer products (ker_openacc_compute_fermion_force,
 * direct_product_of_fermions_into_auxmat): nmu = 4 (default) or 2 links of a
 * site per thread, each vector of the thread's own site read once per shift;
 * 0 one thread per link.  Results are bit-identical.  Returns the previous
 * setting (-1 and a recorded error for another value). */
int ostp_set_direct_product_kernel(int nmu);

/* ===================== fermion measurements (Meas/ferm_meas.c) ========== */
/* Z2-noise chiral condensate, ferm_meas.c:282-320 for nvec noise vectors:
 * noise k is ostp_generate_vec3_soa_z2noise(seed, stream 2k) on even sites and
 * (seed, 2k+1) on odd sites; M^-1 eta by eo_inversion's even-odd split with
 * the even systems solved together by the batched solver (half_inner: FP16
 * storage, else FP32; FP64 reliable updates and true residual); cond[2k],
 * cond[2k+1] = Re, Im of (<eta_e, x_e> + <eta_o, x_o>) degeneracy 0.25 / V.
 * cg_return[k] may be NULL.  Returns the number of converged solves.  No
 * single reference counterpart (fermion_measures does this per vector and
 * writes it to its output file). */
int ostp_chiral_condensate(const su3_soa *u, const su3_soa_f *u_f, ferm_param *pars, int nvec, uint64_t seed,
                           double degeneracy, double res, int max_cg, int half_inner, double *cond,
                           int *cg_return);

/* ============================== eigenvalue bounds (OpenAcc/find_min_max.h) */
/* find_min_max.h:9-14: power iterations of M = m^2 - Deo Doe from loc_p (a
 * Gaussian start vector, overwritten), until the norm changes by <= 1e-5
 * relatively; returns the last norm.  Device-resident loop. */
double ker_find_max_eigenvalue_openacc(const su3_soa *u, ferm_param *pars, vec3_soa *loc_r, vec3_soa *loc_h,
                                       vec3_soa *loc_p);
/* find_min_max.h:17-23: the same on (m^2 + (max - m^2) - m^2) - Deo Doe,
 * returns max - norm (the reference's formula, left unused by its caller) */
double ker_find_min_eigenvalue_openacc(const su3_soa *u, ferm_param *pars, vec3_soa *loc_r, vec3_soa *loc_h,
                                       vec3_soa *loc_p, double max);
/* find_min_max.h:26-32: minmax[0] = m^2, minmax[1] = max eigenvalue (loc_p1) */
void find_min_max_eigenvalue_soloopenacc(const su3_soa *u, ferm_param *pars, vec3_soa *loc_r, vec3_soa *loc_h,
                                         vec3_soa *loc_p1, vec3_soa *loc_p2, double *minmax);
/* The first two with the iteration count (no reference counterpart). */
double ostp_find_max_eigenvalue(const su3_soa *u, ferm_param *pars, vec3_soa *loc_r, vec3_soa *loc_h,
                                vec3_soa *loc_p, int *iterations);
double ostp_find_min_eigenvalue(const su3_soa *u, ferm_param *pars, vec3_soa *loc_r, vec3_soa *loc_h,
                                vec3_soa *loc_p, double max, int *iterations);

/* ====================== rational approximations (RationalApprox/rationalapprox.h) */
/* Host only.  rationalapprox.h:18-20: file names (malloc'ed, caller frees) */
char *rational_approx_filename(double error, int exponent_num, int exponent_den, double lambda_min);
char *rational_approx_filename_old(int approx_order, int exponent_num, int exponent_den, double lambda_min);
/* rationalapprox.h:23-24: read a .REMEZ file (0 ok, 1 missing, 2 malformed) */
int rationalapprox_read(RationalApprox *rational_approx);
int rationalapprox_read_custom_nomefile(RationalApprox *rational_approx, char *nomefile);
/* rationalapprox.h:27 */
void rationalapprox_save(const char *nomefile, RationalApprox *rational_approx);
/* rationalapprox.h:29-33: rescale to [0.95 min, 1.05 max]; renormalise to
 * lambda_max = 1; evaluate a0 + sum_k a_k / (x + b_k) */
void rescale_rational_approximation(RationalApprox *in, RationalApprox *out, double *minmax);
void renormalize_rational_approximation(RationalApprox *in, RationalApprox *out);
double rational_approx_evaluate(RationalApprox *in, double x);

/* ================================== learned preconditioner (new) ======= */
/* No reference counterpart.  A gauge-equivariant network of two
 * parallel-transport layers with OSTP_PTC_CHANNELS channels approximating
 * A^-1 for A = m^2 - Deo Doe on even sites:
 *   z = N r = w0 r + sum_c B_c^dag B_c r,  B_c = sum_nu w[c][nu] T_nu,
 * T_nu the 8 parallel transports of D_oe (forward nu = 0..3, backward 4..7).
 * N is Hermitian, positive for w0 > 0.  Parameters: params[0] = w0,
 * params[1 + 2(8c + nu)] = Re w[c][nu], params[2 + 2(8c + nu)] = Im w[c][nu].
 * FP32 links/phases (the FP32 copies of the solver), BF16 channel mixing on
 * the matrix cores with FP32 accumulation; FP64 input and output vectors. */
#define OSTP_PTC_CHANNELS 12
#define OSTP_PTC_NPARAM (1 + 2 * 8 * OSTP_PTC_CHANNELS)
int ostp_ptc_apply(const float *params, const su3_soa_f *u_f, const float_soa *phases_f, const vec3_soa *in,
                   vec3_soa *out);
/* Training on a pair (b, x = A^-1 b): returns L = |e|^2 (a_norm = 0) or
 * L = <e, A e> (a_norm = 1, the error norm CG minimises; needs the FP64
 * links u and pars->phases), e = N b - x, and writes dL/dparams
 * (OSTP_PTC_NPARAM doubles, same order as params; Re/Im derivatives
 * separately).  pars->ferm_mass gives A; pars->phases_f the FP32 phases. */
double ostp_ptc_loss_grad(const float *params, const su3_soa *u, const su3_soa_f *u_f, ferm_param *pars,
                          const vec3_soa *b, const vec3_soa *x, int a_norm, double *grad);
/* Flexible CG (Polak-Ribiere beta, device-resident scalars) on
 * m^2 - Deo Doe preconditioned by N; the exit test and the final check use
 * the FP64 residual as ker_invert_openacc does (SAFETY 0.95, success iff
 * the true residual <= res).  If flexible CG ends above res, plain CG
 * (ker_invert_openacc semantics) continues from its iterate.  *cg_return =
 * flexible-CG iterations, *fallback_cg = plain-CG iterations (may be NULL),
 * *true_res = final ||b - A x|| / ||b|| (may be NULL).  u_f: FP32 copy of u;
 * pars->phases_f: FP32 phases. */
int ostp_fcg_solve(const su3_soa *u, ferm_param *pars, const su3_soa_f *u_f, const float *params,
                   vec3_soa *solution, const vec3_soa *in, double res, int max_cg, int *cg_return,
                   int *fallback_cg, double *true_res);
/* The same flexible CG run in FP32 (FP32 links, phases and vectors, N
 * applied to FP32 fields) with FP64 reliable updates: whenever the FP32
 * residual has dropped by inner_res since the last update, the accumulated
 * FP32 update is added to x in FP64 and the FP64 true residual b - A x
 * replaces the FP32 one, keeping the search direction (0.1: an update at
 * every 10x drop).  The exit test is the FP64 true residual
 * ||b - A x|| / ||b|| <= res.  An update that leaves more than
 * (1 + inner_res) / 2 of the previous FP64 residual, or an exhausted budget,
 * hands over to plain CG as in ostp_fcg_solve.  *cg_return = FP32 iterations, *outer_return = reliable
 * updates (may be NULL).  It has no counterpart in the reference (whose mixed
 * solver, OpenAcc/inverter_mixed_precision.c, is unpreconditioned). */
int ostp_fcg_mixed_solve(const su3_soa *u, ferm_param *pars, const su3_soa_f *u_f, const float *params,
                         vec3_soa *solution, const vec3_soa *in, double res, int max_cg, double inner_res,
                         int *cg_return, int *outer_return, int *fallback_cg, double *true_res);

/* ======================================================== generators ===== */
/* Reference field rules (OpenAcc/random_assignement.c:67-92,106-142,188-250)
 * drawn from a counter-based Philox4x32-10 stream keyed by `seed` instead of
 * the sequential dSFMT state, so every site is generated independently on the
 * device.  `stream` separates independent sources under one seed. */
void ostp_generate_vec3_soa_gauss(vec3_soa *vect, uint64_t seed, uint32_t stream);
/* generate_Momenta_gauss (random_assignement.c:145-185): the reference's
 * construction of the 8 link momenta from eight variance-1/2 normals, drawn
 * from the same counter-based stream as the other generators */
void ostp_generate_Momenta_gauss(thmat_soa *mom8, uint64_t seed, uint32_t stream);
void ostp_generate_vec3_soa_z2noise(vec3_soa *vect, uint64_t seed, uint32_t stream);
void ostp_generate_Conf_cold(su3_soa *conf, double factor, uint64_t seed);
/* U = exp(i*beta*H), H traceless Hermitian with standard-normal entries */
void ostp_generate_Conf_expH(su3_soa *conf, double beta, uint64_t seed);

/* ======================================================== timing (new) === */
/* Mean device time in microseconds of one Deo (eo=0) or Doe (eo=1) launch
 * (FP64 when single==0, FP32 otherwise), over `reps` back-to-back launches,
 * timed with HIP events on the library's stream. */
double ostp_time_dslash(const void *u, void *out, const void *in, const void *phases,
                        int eo, int single, int reps);
/* Mean device time in microseconds of one streaming pass over `bytes` bytes:
 * kind 0 copies src -> dst, 1 reads src, 2 writes dst; 128-bit non-temporal
 * accesses, `unroll` (1, 2, 4, 8) loads in flight per thread, `grid`
 * workgroups of 256 threads.  bytes must be a multiple of 4096 x unroll.
 * The HBM ceiling the Dslash bandwidth is compared with. */
double ostp_time_stream(int kind, void *dst, const void *src, size_t bytes, int reps, int grid,
                        int unroll);
/* HIP events on the library's stream: start, then (after queueing work)
 * stop, then the elapsed device time in microseconds (waits for stop). */
int ostp_timer_start(void);
int ostp_timer_stop(void);
double ostp_timer_elapsed_us(void);
/* Alternate the walk direction of consecutive Dslash launches (on by
 * default): a launch then starts on the links the previous launch streamed
 * last, which the memory-side cache still holds.  Results do not depend on
 * it (reductions are summed in lattice-tile order).  Returns the previous
 * setting. */
int ostp_set_dslash_serpentine(int on);

#ifdef __cplusplus
}
#endif
#endif /* OPENSTAPLE_AMD_H */
