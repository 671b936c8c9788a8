/*
 * CPU oracle for the OpenStaPLE fermion hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference loops, used by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the checker.
 * The product library (openstaple_amd/) never links or calls it.
 *
 * Memory layout is byte-identical to the reference structure-of-arrays types
 * (OpenAcc/struct_c_def.h:16-58): a vec3_soa is three planes of `sizeh`
 * interleaved (re,im) doubles; an su3_soa is three rows of vec3_soa, plus a
 * double plane K when parallel tempering is compiled in; a gauge field is
 * su3_soa[8] indexed 2*mu+parity; phases are double_soa[8].  The only change
 * is that the lattice extents are runtime values held in or_geom.
 *
 * Parity status of each function is documented next to its declaration and in
 * the tests that pin it.
 */
#ifndef OSTP_ORACLE_H
#define OSTP_ORACLE_H

#include <complex.h>
#include <stddef.h>

typedef double complex or_dc;
typedef float complex or_fc;

typedef struct {
    int nd[4];      /* nd0..nd3 (d0 fastest) */
    long sizeh;     /* nd0*nd1*nd2*nd3/2 */
    int par_temp;   /* su3_soa carries the K plane */
} or_geom;

/* ---- geometry (Mpi/geometry_multidev.h:219-229, OpenAcc/geometry.c:14-100) */
void or_geom_init(or_geom *g, int nd0, int nd1, int nd2, int nd3, int par_temp);
size_t or_vec3_bytes(const or_geom *g);
size_t or_su3_bytes(const or_geom *g);
int  or_snum(const or_geom *g, int d0, int d1, int d2, int d3);
/* nnp/nnm laid out [sizeh][4][2] as nnp_openacc/nnm_openacc (geometry.h:59-60) */
void or_nn_tables(const or_geom *g, int *nnp, int *nnm);

/* ---- U(1) phases (OpenAcc/backfield.c:20-187); map = {xmap,ymap,zmap,tmap},
 *      bf = {ex,ey,ez,bx,by,bz}; phases = double[8*sizeh]                   */
void or_calc_u1_phases(const or_geom *g, const int map[4], const double bf[6],
                       double im_chem_pot, double ferm_charge, double *phases);
void or_phases_to_float(const or_geom *g, const double *phases, float *phases_f);

/* ---- staggered Dslash (OpenAcc/fermion_matrix.c:47-157, matvecmul.h:88-172) */
void or_acc_Deo(const or_geom *g, const void *u, void *out, const void *in, const double *phases);
void or_acc_Doe(const or_geom *g, const void *u, void *out, const void *in, const double *phases);
/* FP32 twins (what the generated sp_ code computes: float data, cosf/sinf) */
void or_acc_Deo_f(const or_geom *g, const void *u_f, void *out, const void *in, const float *phases_f);
void or_acc_Doe_f(const or_geom *g, const void *u_f, void *out, const void *in, const float *phases_f);

/* out = (m^2+shift) in - Deo Doe in ; temp holds Doe in (fermion_matrix.c:723-746) */
void or_fermion_matrix_multiplication_shifted(const or_geom *g, const void *u, void *out,
        const void *in, void *temp, const double *phases, double mass, double shift);
void or_fermion_matrix_multiplication_shifted_f(const or_geom *g, const void *u_f, void *out,
        const void *in, void *temp, const float *phases_f, double mass, double shift);

/* ---- BLAS (OpenAcc/fermionic_utilities.c:126-313) */
double or_l2norm2(const or_geom *g, const void *v);
double or_real_scal_prod(const or_geom *g, const void *a, const void *b);
or_dc  or_scal_prod(const or_geom *g, const void *a, const void *b);
double or_l2norm2_f(const or_geom *g, const void *v);
double or_real_scal_prod_f(const or_geom *g, const void *a, const void *b);

/* ---- solvers */
typedef struct {
    int restarting_every;          /* inverter_tricks.restartingEvery */
    double mixed_precision_delta;  /* inverter_tricks.mixedPrecisionDelta */
} or_tricks;

/* ker_invert_openacc (OpenAcc/inverter_full.c:19-132); returns 1 success / 0 failure */
int or_ker_invert(const or_geom *g, const void *u, const double *phases, double mass,
                  void *solution, const void *in, double res,
                  void *r, void *h, void *s, void *p, int max_cg, double shift,
                  const or_tricks *tr, int *cg_return);

/* inverter_mixed_precision (OpenAcc/inverter_mixedp.c:41-181) */
int or_inverter_mixed_precision(const or_geom *g, const void *u, const void *u_f,
                  const double *phases, const float *phases_f, double mass,
                  void *solution, const void *in, double res, int max_cg, double shift,
                  const or_tricks *tr, int *cg_return);

/* multishift_invert (OpenAcc/inverter_multishift_full.c:23-252) */
int or_multishift_invert(const or_geom *g, const void *u, const double *phases, double mass,
                  int order, const double *RA_b, void *out, const void *in, double residuo,
                  void *r, void *h, void *s, void *p, void *shiftferm, int max_cg,
                  int *cg_return);
/* recombine_shifted_vec3_to_vec3 (inverter_multishift_full.c:254-282) */
void or_recombine(const or_geom *g, const void *in_shifted, const void *in, void *out,
                  int order, double a0, const double *RA_a);

/* precision conversions (OpenAcc/float_double_conv.h:14-56) */
void or_vec3_d2f(const or_geom *g, const void *d, void *f);
void or_vec3_f2d(const or_geom *g, const void *f, void *d);
void or_su3_d2f(const or_geom *g, const void *d, void *f, int nlinks);

/* ---- generators (counter-based Philox4x32-10; the library restates them on device) */
void or_philox4x32_10(const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4]);
void or_gen_vec3_gauss(const or_geom *g, void *v, unsigned long long seed, unsigned int stream);
void or_gen_vec3_z2(const or_geom *g, void *v, unsigned long long seed, unsigned int stream);
void or_gen_conf_cold(const or_geom *g, void *u, double eps, unsigned long long seed);
void or_gen_conf_expH(const or_geom *g, void *u, double beta, unsigned long long seed);

/* ---- gauge part (gauge.c): stout smearing, its force chain, unitarisation,
 *      fermion-force building blocks.  tamat/thmat fields are 8 x 64*sizeh bytes. */
void or_calc_loc_staples(const or_geom *g, const void *u, void *stap, double c_zero);
void or_rho_times_conf_times_staples_ta(const or_geom *g, const void *u, const void *stap, void *qa,
                                        double rho, double c_zero);
void or_exp_minus_QA_times_conf(const or_geom *g, const void *u, const void *qa, void *uout);
void or_stout_isotropic(const or_geom *g, const void *u, void *uout, double rho, double c_zero);
void or_compute_lambda(const or_geom *g, void *L, const void *SP, const void *U, const void *QA);
void or_compute_sigma(const or_geom *g, const void *L, const void *U, void *S, const void *QA, double rho);
void or_sigma_prime_to_sigma(const or_geom *g, void *Sigma, void *Lambda, void *QA, const void *U, double rho,
                             double c_zero);
void or_unitarize_conf(const or_geom *g, void *u);
void or_direct_product_into_auxmat(const or_geom *g, const void *s, const void *h, void *aux, double a);
void or_multiply_backfield_times_force(const or_geom *g, const double *phases, const void *aux, void *out);
void or_conf_times_force_ta(const or_geom *g, const void *u, const void *aux, void *ipdot);

/* ---- rational approximations (RationalApprox/rationalapprox.c:83-236) */
#define OR_MAX_APPROX_ORDER 25
typedef struct {
    int exponent_num, exponent_den, approx_order;
    double lambda_min, lambda_max;
    int gmp_remez_precision;
    double error, RA_a0;
    double RA_a[OR_MAX_APPROX_ORDER], RA_b[OR_MAX_APPROX_ORDER];
} or_ratapprox;
int or_ratapprox_read(or_ratapprox *ra, const char *filename);
double or_ratapprox_evaluate(const or_ratapprox *ra, double x);
void or_ratapprox_rescale(const or_ratapprox *in, or_ratapprox *out, const double minmax[2]);

/* eo_inversion (Meas/ferm_meas.c:50-71): full-matrix solve through the
 * even-odd CG; returns the CG status */
/* power method of find_min_max.c:21-97 (loc_p: start vector, overwritten) */
double or_find_max_eigenvalue(const or_geom *g, const void *u, const double *phases, double mass, void *loc_r,
                              void *loc_h, void *loc_p, int *iters);
double or_find_min_eigenvalue(const or_geom *g, const void *u, const double *phases, double mass, void *loc_r,
                              void *loc_h, void *loc_p, double max, int *iters);
int or_eo_inversion(const or_geom *g, const void *u, const double *phases, double mass, double res, int max_cg,
                    const void *in_e, const void *in_o, void *out_e, void *out_o, void *phi_e, void *phi_o,
                    void *r, void *h, void *s, void *p, const or_tricks *tr, int *cg_return);

/* ---- parallel tempering (OpenAcc/HPT_utilities.c), single rank */
/* init_k (:30-310); conf = su3_soa[8] with K planes (par_temp geometry), map =
 * {xmap,ymap,zmap,tmap}; def_out = 100 ints in the GAUGE_ACT_TLSM layout of
 * defect_info (min[4][4], max[4][4], min_TLSM[2][4][4], max_TLSM[2][4][4],
 * def_axis_mapped, def_mapped_perp_dir[3]), written when defect_info_config == 0 */
void or_init_k(const or_geom *g, const int map[4], void *conf, double c_r, int def_axis, const int def_vec[3],
               int *def_out, int defect_info_config);
int or_metro_swap(double delta_s, double u);
typedef void (*or_pt_action_fn)(const double *cr_of_replica, const int *label, double *S_of_label, void *user);
typedef double (*or_pt_uniform_fn)(void *user);
void or_manage_replica_swaps(int n, int is_evenodd, const double *cr_vec, int *label, or_pt_action_fn action,
                             or_pt_uniform_fn uniform, void *user, int *swap_num, int *all_swap_vet,
                             int *acceptance_vet);

/* ---- parallel-transport preconditioner (oracle/ptc.c): w = C x 8 complex, row c */
void or_ptc_transports(const or_geom *g, const void *u, const double *ph, const void *in, void *out);
void or_ptc_adjoint_add(const or_geom *g, const void *u, const double *ph, const void *in, void *out);
void or_ptc_apply(const or_geom *g, const void *u, const double *ph, int C, double w0, const or_dc *w,
                  const void *r, void *z);
double or_ptc_loss_grad(const or_geom *g, const void *u, const double *ph, int C, double w0, const or_dc *w,
                        const void *b, const void *x, double *grad);

#endif
