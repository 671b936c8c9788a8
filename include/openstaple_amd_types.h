/*
 * openstaple_amd_types.h -- complete definitions of the reference's by-value
 * and field-access types, for callers that do not compile against the
 * reference headers (tests, Python ctypes, new C code).  A host built on the
 * reference headers must NOT include this file: its own definitions are the
 * same layouts.
 *
 * The structure-of-arrays field types (vec3_soa, su3_soa, ...) have a size that
 * depends on the lattice; they stay incomplete here and are addressed through
 * byte offsets: plane k of a vec3_soa starts at k*16*sizeh, row r / column c
 * of an su3_soa at (3r+c)*16*sizeh, and consecutive su3_soa are
 * (144 + 8*par_temp)*sizeh bytes apart (FP32 twins: half of each).
 */
#ifndef OPENSTAPLE_AMD_TYPES_H
#define OPENSTAPLE_AMD_TYPES_H

#include "openstaple_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MAX_APPROX_ORDER 25 /* rationalapprox.h:8 */

struct RationalApprox_t {   /* RationalApprox/rationalapprox.h:15-26 */
    int exponent_num;
    int exponent_den;
    int approx_order;
    double lambda_min;
    double lambda_max;
    int gmp_remez_precision;
    double error;
    double RA_a0;
    double RA_a[MAX_APPROX_ORDER];
    double RA_b[MAX_APPROX_ORDER];
};

struct ferm_param_t {       /* Include/fermion_parameters.h:9-41 */
    double ferm_mass;
    int degeneracy;
    int number_of_ps;
    char name[30];
    double ferm_charge;
    double ferm_im_chem_pot;
    int index_of_the_first_ps;
    int index_of_the_first_shift;
    double_soa *phases;
    double_soa *mag_re;
    double_soa *mag_im;
    int printed_bf_dbg_info;
    float_soa *phases_f;
    RationalApprox approx_fi_mother;
    RationalApprox approx_md_mother;
    RationalApprox approx_li_mother;
    RationalApprox approx_fi;
    RationalApprox approx_md;
    RationalApprox approx_li;
};

struct inverter_package_t { /* OpenAcc/inverter_package.h:12-29 */
    const su3_soa *u;
    const su3_soa_f *u_f;
    vec3_soa *ferm_shift_temp;
    vec3_soa_f *ferm_shift_temp_f;
    int nshifts;
    vec3_soa *loc_r;
    vec3_soa *loc_h;
    vec3_soa *loc_s;
    vec3_soa *loc_p;
    vec3_soa_f *loc_r_f;
    vec3_soa_f *loc_h_f;
    vec3_soa_f *loc_s_f;
    vec3_soa_f *loc_p_f;
    vec3_soa_f *out_f;
};

struct md_param_t {         /* OpenAcc/md_parameters.h:6-19 */
    int no_md;
    int gauge_scale;
    double t;
    double residue_metro;
    double expected_max_eigenvalue;
    int singlePrecMD;
    double residue_md;
    int max_cg_iterations;
    int recycleInvsForce;
    int extrapolateInvsForce;
};

struct action_param_t {     /* OpenAcc/action.h:6-18 */
    double beta;
    int stout_steps;
    double stout_rho;
    int topo_action;
    double barrier;
    double width;
    char topo_file_path[20];
    int topo_stout_steps;
    double topo_rho;
};

struct alloc_settings_t {   /* OpenAcc/alloc_settings.h:4-20 */
    int NDiffFlavs;
    int NPS_tot;
    int conf_acc_size;
    int maxNeededShifts;
    int maxApproxOrder;
    int revTestAllocations;
    int diagnosticsAllocations;
    int singlePrecCoreAllocations;
    int singlePrecExtendedAllocations;
    int stoutAllocations;
    int num_replicas;
};

struct bf_param_t {         /* OpenAcc/backfield_parameters.h:5-9 */
    double ex, ey, ez, bx, by, bz;
};

struct inv_tricks_t {       /* Include/inverter_tricks.h:4-11 */
    int singlePInvAccelMultiInv;
    int useMixedPrecision;
    double mixedPrecisionDelta;
    int restartingEvery;
};

extern inv_tricks inverter_tricks; /* Include/inverter_tricks.c:6 */

struct rep_info_t {         /* Include/rep_info.h:4-11 */
    int replicas_total_number;
    int is_evenodd;
    int defect_boundary;
    int defect_coordinates[3];
    double *cr_vec;
    int *label;
};

/* OpenAcc/HPT_utilities.h:14-32 as compiled with GAUGE_ACT_TLSM (the
 * tree-level Symanzik action, the library default).  With the Wilson action
 * (ostp_set_stout(..., tlsm = 0)) the reference struct has no *_TLSM members;
 * init_k then writes the ostp_defect_info_wilson layout below. */
struct defect_info_t {
    int defect_swap_min[4][4];
    int defect_swap_max[4][4];
    int defect_swap_min_TLSM[2][4][4];
    int defect_swap_max_TLSM[2][4][4];
    int def_axis_mapped;
    int def_mapped_perp_dir[3];
};
typedef struct ostp_defect_info_wilson_t {
    int defect_swap_min[4][4];
    int defect_swap_max[4][4];
    int def_axis_mapped;
    int def_mapped_perp_dir[3];
} ostp_defect_info_wilson;

#ifdef __cplusplus
}
#endif
#endif /* OPENSTAPLE_AMD_TYPES_H */
