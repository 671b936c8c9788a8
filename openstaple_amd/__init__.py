"""Host-side mirror of the OpenStaPLE fermion hot path on MI355X.

The product is the C-ABI shared library ``libopenstaple_amd.so`` (HIP kernels
for gfx950, declared in ``include/openstaple_amd.h``).  This module binds it
with ctypes and provides reference-layout host containers so Python callers
(tests, benchmarks) drive the same entry points a C host would.  There is no
CPU fallback: importing raises if the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libopenstaple_amd.so"

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is missing: build the HIP library first "
        "(python -m openstaple_amd.build, or __graft_entry__.build())"
    )

lib = C.CDLL(str(LIB_PATH))

MAX_APPROX_ORDER = 25
INVERTER_SUCCESS = 1
INVERTER_FAILURE = 0


# ------------------------------------------------------------ reference types
class RationalApprox(C.Structure):  # RationalApprox/rationalapprox.h:15-26
    _fields_ = [
        ("exponent_num", C.c_int),
        ("exponent_den", C.c_int),
        ("approx_order", C.c_int),
        ("lambda_min", C.c_double),
        ("lambda_max", C.c_double),
        ("gmp_remez_precision", C.c_int),
        ("error", C.c_double),
        ("RA_a0", C.c_double),
        ("RA_a", C.c_double * MAX_APPROX_ORDER),
        ("RA_b", C.c_double * MAX_APPROX_ORDER),
    ]


class FermParam(C.Structure):  # Include/fermion_parameters.h:9-41
    _fields_ = [
        ("ferm_mass", C.c_double),
        ("degeneracy", C.c_int),
        ("number_of_ps", C.c_int),
        ("name", C.c_char * 30),
        ("ferm_charge", C.c_double),
        ("ferm_im_chem_pot", C.c_double),
        ("index_of_the_first_ps", C.c_int),
        ("index_of_the_first_shift", C.c_int),
        ("phases", C.c_void_p),
        ("mag_re", C.c_void_p),
        ("mag_im", C.c_void_p),
        ("printed_bf_dbg_info", C.c_int),
        ("phases_f", C.c_void_p),
        ("approx_fi_mother", RationalApprox),
        ("approx_md_mother", RationalApprox),
        ("approx_li_mother", RationalApprox),
        ("approx_fi", RationalApprox),
        ("approx_md", RationalApprox),
        ("approx_li", RationalApprox),
    ]


class InverterPackage(C.Structure):  # OpenAcc/inverter_package.h:12-29
    _fields_ = [
        ("u", C.c_void_p),
        ("u_f", C.c_void_p),
        ("ferm_shift_temp", C.c_void_p),
        ("ferm_shift_temp_f", C.c_void_p),
        ("nshifts", C.c_int),
        ("loc_r", C.c_void_p),
        ("loc_h", C.c_void_p),
        ("loc_s", C.c_void_p),
        ("loc_p", C.c_void_p),
        ("loc_r_f", C.c_void_p),
        ("loc_h_f", C.c_void_p),
        ("loc_s_f", C.c_void_p),
        ("loc_p_f", C.c_void_p),
        ("out_f", C.c_void_p),
    ]


class MdParam(C.Structure):  # OpenAcc/md_parameters.h:6-19
    _fields_ = [
        ("no_md", C.c_int),
        ("gauge_scale", C.c_int),
        ("t", C.c_double),
        ("residue_metro", C.c_double),
        ("expected_max_eigenvalue", C.c_double),
        ("singlePrecMD", C.c_int),
        ("residue_md", C.c_double),
        ("max_cg_iterations", C.c_int),
        ("recycleInvsForce", C.c_int),
        ("extrapolateInvsForce", C.c_int),
    ]


class ActionParam(C.Structure):  # OpenAcc/action.h:6-18
    _fields_ = [
        ("beta", C.c_double),
        ("stout_steps", C.c_int),
        ("stout_rho", C.c_double),
        ("topo_action", C.c_int),
        ("barrier", C.c_double),
        ("width", C.c_double),
        ("topo_file_path", C.c_char * 20),
        ("topo_stout_steps", C.c_int),
        ("topo_rho", C.c_double),
    ]


class AllocSettings(C.Structure):  # OpenAcc/alloc_settings.h:4-20
    _fields_ = [(n, C.c_int) for n in (
        "NDiffFlavs", "NPS_tot", "conf_acc_size", "maxNeededShifts", "maxApproxOrder", "revTestAllocations",
        "diagnosticsAllocations", "singlePrecCoreAllocations", "singlePrecExtendedAllocations", "stoutAllocations",
        "num_replicas")]


class RhmcInfo(C.Structure):  # ostp_rhmc_info
    _fields_ = [
        ("iteration", C.c_int),
        ("accepted", C.c_int),
        ("md_multishift_iterations", C.c_int),
        ("action_in", C.c_double),
        ("action_mom_in", C.c_double),
        ("action_ferm_in", C.c_double),
        ("action_fin", C.c_double),
        ("action_mom_fin", C.c_double),
        ("action_ferm_fin", C.c_double),
        ("delta_S", C.c_double),
        ("p1", C.c_double),
        ("p2", C.c_double),
    ]


UNIFORM_FN = C.CFUNCTYPE(C.c_double, C.c_void_p)


class BfParam(C.Structure):  # OpenAcc/backfield_parameters.h:5-9
    _fields_ = [(n, C.c_double) for n in ("ex", "ey", "ez", "bx", "by", "bz")]


class BatchSystem(C.Structure):  # ostp_batch_system
    _fields_ = [
        ("nd", C.c_int * 4),
        ("par_temp", C.c_int),
        ("u", C.c_void_p),
        ("u_f", C.c_void_p),
        ("phases", C.c_void_p),
        ("phases_f", C.c_void_p),
        ("in_", C.c_void_p),
        ("out", C.c_void_p),
        ("mass", C.c_double),
    ]


class RepInfo(C.Structure):  # Include/rep_info.h:4-11
    _fields_ = [
        ("replicas_total_number", C.c_int),
        ("is_evenodd", C.c_int),
        ("defect_boundary", C.c_int),
        ("defect_coordinates", C.c_int * 3),
        ("cr_vec", C.POINTER(C.c_double)),
        ("label", C.POINTER(C.c_int)),
    ]


class DefectInfo(C.Structure):  # OpenAcc/HPT_utilities.h:14-32 (GAUGE_ACT_TLSM layout)
    _fields_ = [
        ("defect_swap_min", (C.c_int * 4) * 4),
        ("defect_swap_max", (C.c_int * 4) * 4),
        ("defect_swap_min_TLSM", ((C.c_int * 4) * 4) * 2),
        ("defect_swap_max_TLSM", ((C.c_int * 4) * 4) * 2),
        ("def_axis_mapped", C.c_int),
        ("def_mapped_perp_dir", C.c_int * 3),
    ]


class DefectInfoWilson(C.Structure):  # the same struct compiled without GAUGE_ACT_TLSM
    _fields_ = [
        ("defect_swap_min", (C.c_int * 4) * 4),
        ("defect_swap_max", (C.c_int * 4) * 4),
        ("def_axis_mapped", C.c_int),
        ("def_mapped_perp_dir", C.c_int * 3),
    ]


PTC_CHANNELS = 12
PTC_NPARAM = 1 + 2 * 8 * PTC_CHANNELS

PT_ACTION_FN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_void_p)
PT_UNIFORM_FN = C.CFUNCTYPE(C.c_double, C.c_void_p)
# io.c:553-559 callback: (idxh machine, parity, dirmachine, data, endianness disagreement, FILE*)
RW_SITE_FN = C.CFUNCTYPE(None, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p)


class DComplex(C.Structure):
    """A C99 double complex return value: on x86-64 it travels in two SSE
    registers, exactly like this two-double struct."""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]

    def __complex__(self):
        return complex(self.re, self.im)

assert C.sizeof(RationalApprox) == 456
assert FermParam.phases.offset == 72 and FermParam.phases_f.offset == 104
assert C.sizeof(InverterPackage) == 112
assert C.sizeof(BatchSystem) == 80
assert C.sizeof(RepInfo) == 40 and C.sizeof(DefectInfo) == 400 and C.sizeof(DefectInfoWilson) == 144

# ------------------------------------------------------------------ signatures
_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_SIGS = {
    "ostp_init": (_I, [_I, _I, _I, _I, _I]),
    "ostp_finalize": (None, []),
    "ostp_set_direction_map": (_I, [_I, _I, _I, _I]),
    "ostp_check_sizes": (_I, [C.c_size_t, C.c_size_t, C.c_size_t]),
    "ostp_sizeh": (C.c_long, []),
    "ostp_acc_create": (_I, [_P, C.c_size_t]),
    "ostp_acc_copyin": (_I, [_P, C.c_size_t]),
    "ostp_acc_update_device": (_I, [_P, C.c_size_t]),
    "ostp_acc_update_self": (_I, [_P, C.c_size_t]),
    "ostp_acc_delete": (_I, [_P]),
    "ostp_acc_deviceptr": (_P, [_P]),
    "ostp_acc_is_present": (_I, [_P, C.c_size_t]),
    "ostp_malloc": (_P, [C.c_size_t]),
    "ostp_free": (None, [_P]),
    "ostp_memcpy_htod": (_I, [_P, _P, C.c_size_t]),
    "ostp_memcpy_dtoh": (_I, [_P, _P, C.c_size_t]),
    "ostp_memcpy_dtod": (_I, [_P, _P, C.c_size_t]),
    "ostp_memset": (_I, [_P, _I, C.c_size_t]),
    "ostp_synchronize": (_I, []),
    "ostp_set_abort_on_error": (None, [_I]),
    "ostp_last_error": (C.c_char_p, []),
    "ostp_clear_error": (None, []),
    "ostp_set_sync_at_return": (_I, [_I]),
    "ostp_set_inverter_tricks": (None, [_I, _I, _D, _I]),
    "ostp_compute_nnp_and_nnm": (_I, [_P, _P]),
    "calc_u1_phases": (None, [_P, BfParam, _D, _D]),
    "ostp_phases_to_float": (None, [_P, _P]),
    "acc_Deo": (None, [_P, _P, _P, _P]),
    "acc_Doe": (None, [_P, _P, _P, _P]),
    "acc_Deo_f": (None, [_P, _P, _P, _P]),
    "acc_Doe_f": (None, [_P, _P, _P, _P]),
    "fermion_matrix_multiplication": (None, [_P, _P, _P, _P, _P]),
    "fermion_matrix_multiplication_shifted": (None, [_P, _P, _P, _P, _P, _D]),
    "fermion_matrix_multiplication_shifted_f": (None, [_P, _P, _P, _P, _P, _D]),
    "l2norm2_global": (_D, [_P]),
    "real_scal_prod_global": (_D, [_P, _P]),
    "ostp_scal_prod_global": (None, [_P, _P, C.POINTER(_D), C.POINTER(_D)]),
    "scal_prod_global": (DComplex, [_P, _P]),
    "rw_iterate_on_global_sites_lx_xyzt_axis_ordering": (None, [RW_SITE_FN, _P, _P, _I]),
    "combine_in1xfactor_plus_in2": (None, [_P, _D, _P, _P]),
    "combine_add_factor_x_in2_to_in1": (None, [_P, _P, _D]),
    "combine_in1xferm_mass_minus_in2": (None, [_P, _D, _P]),
    "combine_in1_minus_in2": (None, [_P, _P, _P]),
    "assign_in_to_out": (None, [_P, _P]),
    "set_vec3_soa_to_zero": (None, [_P]),
    "convert_double_to_float_vec3_soa": (None, [_P, _P]),
    "convert_float_to_double_vec3_soa": (None, [_P, _P]),
    "ostp_convert_double_to_float_gauge": (None, [_P, _P]),
    "ker_invert_openacc": (_I, [_P, _P, _P, _P, _D, _P, _P, _P, _P, _I, _D, C.POINTER(_I)]),
    "ker_invert_openacc_noalloc": (_I, [_P, _P, _P, _P, _D, _P, _P, _P, _P, _I, _D, C.POINTER(_I)]),
    "inverter_mixed_precision": (_I, [InverterPackage, _P, _P, _P, _D, _I, _D, C.POINTER(_I)]),
    "multishift_invert": (_I, [_P, _P, _P, _P, _P, _D, _P, _P, _P, _P, _P, _I, C.POINTER(_I)]),
    "recombine_shifted_vec3_to_vec3": (None, [_P, _P, _P, _P]),
    "convergence_messages": (None, [_I, _I]),
    "inverter_wrapper": (_I, [InverterPackage, _P, _P, _P, _D, _I, _D, _I]),
    "inverter_multishift_wrapper": (_I, [InverterPackage, _P, _P, _P, _P, _D, _I, _I]),
    "ostp_multishift_mixed_solve": (_I, [InverterPackage, _P, _P, _P, _P, _D, _I, _I, _P, _P, _P]),
    "ostp_chiral_condensate": (_I, [_P, _P, _P, _I, C.c_uint64, _D, _D, _I, _I, _P, _P]),
    "calc_ipdot_gauge_soloopenacc": (None, [_P, _P, _P]),
    "ostp_time_gauge_force": (_D, [_P, _P, _P, _P, _D, _I]),
    "ostp_set_gauge_force_kernel": (_I, [_I]),
    "ostp_set_direct_product_kernel": (_I, [_I]),
    "ostp_set_lattice_host": (_I, [_I, _I, _I, _I, _I]),
    "ostp_set_multishift_window": (_I, [_I]),
    **{f"{n}{k}": (None, [_P, _P, _P, _P]) for n in ("dM_dmu_eo", "dM_dmu_oe", "d2M_dmu2_eo", "d2M_dmu2_oe")
       for k in range(4)},
    "acc_Deo_wf": (None, [_P, _P, _P, _P, _P, _P]),
    "acc_Doe_wf": (None, [_P, _P, _P, _P, _P, _P]),
    "idphase_dbz": (None, [_P, _P, _P]),
    "ostp_fermion_measures": (_I, [_P, _P, _P, _I, _I, _I, C.c_uint64, _D, _I, _I, _P, C.POINTER(_I)]),
    "print_su3_soa_ASCII": (_I, [_P, C.c_char_p, _I, _P]),
    "read_su3_soa_ASCII": (_I, [_P, C.c_char_p, C.POINTER(_I)]),
    "print_su3_soa_ildg_binary": (_I, [_P, C.c_char_p, _I]),
    "read_su3_soa_ildg_binary": (_I, [_P, C.c_char_p, C.POINTER(_I)]),
    "print_vec3_soa": (None, [_P, C.c_char_p]),
    "read_vec3_soa": (_I, [_P, C.c_char_p]),
    "ostp_save_conf": (_I, [_P, C.c_char_p, _I, _I]),
    "ostp_read_conf": (_I, [_P, C.c_char_p, C.POINTER(_I), _I]),
    "mom_sum_mult": (None, [_P, _P, _P, _I]),
    "mom_exp_times_conf_soloopenacc": (None, [_P, _P, _P, _I]),
    "calc_momenta_action": (_D, [_P, _P, _I]),
    "invert_momenta": (None, [_P]),
    "initialize_md_global_variables": (None, [MdParam]),
    "multistep_2MN_gauge": (None, [_P, _P, _P, _P]),
    "multistep_2MN_SOLOOPENACC": (None, [_P, _P, _P, _P, _P, _P, _I, _P, _P, InverterPackage, _P, _P, _D, _I]),
    "ostp_generate_Momenta_gauss": (None, [_P, C.c_uint64, C.c_uint32]),
    "UPDATE_SOLOACC_UNOSTEP_VERSATILE": (_I, [_P, _D, _D, _I, _I, _I, _I]),
    "ostp_set_rng_seed": (None, [C.c_uint64, C.c_uint32]),
    "ostp_set_metropolis_uniform": (None, [UNIFORM_FN, _P]),
    "ostp_rhmc_last_info": (_I, [C.POINTER(RhmcInfo)]),
    "calc_plaquette_soloopenacc": (_D, [_P, _P, _P]),
    "calc_rettangolo_soloopenacc": (_D, [_P, _P, _P]),
    "ker_find_max_eigenvalue_openacc": (_D, [_P, _P, _P, _P, _P]),
    "ker_find_min_eigenvalue_openacc": (_D, [_P, _P, _P, _P, _P, _D]),
    "find_min_max_eigenvalue_soloopenacc": (None, [_P, _P, _P, _P, _P, _P, _P]),
    "ostp_find_max_eigenvalue": (_D, [_P, _P, _P, _P, _P, _P]),
    "ostp_find_min_eigenvalue": (_D, [_P, _P, _P, _P, _P, _D, _P]),
    "rational_approx_filename": (_P, [_D, _I, _I, _D]),
    "rational_approx_filename_old": (_P, [_I, _I, _I, _D]),
    "rationalapprox_read": (_I, [_P]),
    "rationalapprox_read_custom_nomefile": (_I, [_P, C.c_char_p]),
    "rationalapprox_save": (None, [C.c_char_p, _P]),
    "rescale_rational_approximation": (None, [_P, _P, _P]),
    "renormalize_rational_approximation": (None, [_P, _P]),
    "rational_approx_evaluate": (_D, [_P, _D]),
    "ostp_cg_solve": (_I, [_P, _P, _P, _P, _D, _P, _P, _P, _P, _I, _D, _I, C.POINTER(_I), C.POINTER(_D)]),
    "ostp_mixed_solve": (_I, [InverterPackage, _P, _P, _P, _D, _I, _D, _D, C.POINTER(_I), C.POINTER(_I),
                              C.POINTER(_D)]),
    "ostp_multishift_solve": (_I, [_P, _P, _P, _P, _P, _D, _P, _P, _P, _P, _P, _I, _I, C.POINTER(_I),
                                   C.POINTER(_D)]),
    "ostp_generate_vec3_soa_gauss": (None, [_P, C.c_uint64, C.c_uint32]),
    "ostp_generate_vec3_soa_z2noise": (None, [_P, C.c_uint64, C.c_uint32]),
    "ostp_generate_Conf_cold": (None, [_P, _D, C.c_uint64]),
    "ostp_generate_Conf_expH": (None, [_P, _D, C.c_uint64]),
    "ostp_set_stout": (None, [_D, _I, _D, _I, _I, _I]),
    "stout_wrapper": (None, [_P, _P, _I]),
    "stout_isotropic": (None, [_P, _P, _P, _P, _P, _I]),
    "exp_minus_QA_times_conf": (None, [_P, _P, _P, _P]),
    "compute_lambda": (None, [_P, _P, _P, _P, _P]),
    "compute_sigma": (None, [_P, _P, _P, _P, _P, _I]),
    "compute_sigma_from_sigma_prime_backinto_sigma_prime": (None, [_P, _P, _P, _P, _P, _I]),
    "calc_loc_staples_nnptrick_all_onlyferms": (None, [_P, _P]),
    "RHO_times_conf_times_staples_ta_part": (None, [_P, _P, _P, _I]),
    "unitarize_conf": (None, [_P]),
    "ostp_stout_steps": (_I, []),
    "direct_product_of_fermions_into_auxmat": (None, [_P, _P, _P, _P, _I]),
    "multiply_backfield_times_force": (None, [_P, _P, _P]),
    "multiply_conf_times_force_and_take_ta_nophase": (None, [_P, _P, _P]),
    "ker_openacc_compute_fermion_force": (None, [_P, _P, _P, _P, _P, _P]),
    "fermion_force_soloopenacc": (None, [_P, _P, _P, _P, _P, _I, _P, _D, _P, _P, InverterPackage, _I]),
    "combine_in1_x_fact1_minus_in2_back_into_in2": (None, [_P, _D, _P]),
    "combine_in1_minus_in2_allxfact": (None, [_P, _P, _D, _P]),
    "eo_inversion": (None, [InverterPackage, _P, _D, _I, _P, _P, _P, _P, _P, _P]),
    "init_k": (None, [_P, _D, _I, C.POINTER(_I), _P, _I]),
    "ostp_pt_manage_replica_swaps": (None, [C.POINTER(RepInfo), PT_ACTION_FN, PT_UNIFORM_FN, _P, C.POINTER(_I),
                                            C.POINTER(_I), C.POINTER(_I)]),
    "ostp_pt_init_labels": (None, [C.POINTER(RepInfo)]),
    "ostp_pt_index_of_label0": (_I, [C.POINTER(RepInfo)]),
    "ostp_pt_metropolis": (_I, [_D, _D]),
    "ostp_pt_twin_label": (_I, [_I, _I, _I]),
    "ostp_batched_mixed_solve": (_I, [_I, C.POINTER(BatchSystem), _D, _I, _D, _I, C.POINTER(_I), C.POINTER(_I),
                                      C.POINTER(_D)]),
    "ostp_batched_half_solve": (_I, [_I, C.POINTER(BatchSystem), _D, _I, _D, _I, C.POINTER(_I), C.POINTER(_I),
                                     C.POINTER(_D)]),
    "ostp_set_multishift_batching": (_I, [_I]),
    "ostp_set_batched_graph_iterations": (_I, [_I]),
    "ostp_set_batch_tile": (_I, [_I, _I]),
    "ostp_autotune_report": (_I, [C.c_char_p, _I]),
    "ostp_ptc_apply": (_I, [_P, _P, _P, _P, _P]),
    "ostp_ptc_loss_grad": (_D, [_P, _P, _P, _P, _P, _P, _I, _P]),
    "ostp_fcg_solve": (_I, [_P, _P, _P, _P, _P, _P, _D, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_D)]),
    "ostp_fcg_mixed_solve": (_I, [_P, _P, _P, _P, _P, _P, _D, _I, _D, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I),
                                  C.POINTER(_D)]),
    "ostp_time_dslash": (_D, [_P, _P, _P, _P, _I, _I, _I]),
    "ostp_time_stream": (_D, [_I, _P, _P, C.c_size_t, _I, _I, _I]),
    "ostp_set_dslash_serpentine": (_I, [_I]),
    "ostp_set_multishift_precision": (_I, [_I, _I]),
    "ostp_timer_start": (_I, []),
    "ostp_timer_stop": (_I, []),
    "ostp_timer_elapsed_us": (_D, []),
}

for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED = tuple(_SIGS)


class OstpError(RuntimeError):
    pass


def check(rc: int = 0, what: str = "") -> None:
    """Raise if the library recorded an error (non-aborting mode)."""
    msg = lib.ostp_last_error()
    if msg:
        lib.ostp_clear_error()
        raise OstpError(f"{what}: {msg.decode()}")
    if rc not in (0, None):
        raise OstpError(f"{what}: return code {rc}")


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# ---------------------------------------------------------- lattice helpers
class Lattice:
    """Lattice extents and reference-layout host containers (numpy)."""

    def __init__(self, nd0: int, nd1: int, nd2: int, nd3: int, par_temp: bool = False):
        self.nd = (nd0, nd1, nd2, nd3)
        self.sizeh = nd0 * nd1 * nd2 * nd3 // 2
        self.par_temp = bool(par_temp)

    # byte sizes of the reference structs
    @property
    def vec3_bytes(self) -> int:
        return 48 * self.sizeh

    @property
    def su3_bytes(self) -> int:
        return (144 + (8 if self.par_temp else 0)) * self.sizeh

    @property
    def su3f_bytes(self) -> int:
        return (72 + (4 if self.par_temp else 0)) * self.sizeh

    def init(self) -> None:
        lib.ostp_set_abort_on_error(0)
        rc = lib.ostp_init(*self.nd, int(self.par_temp))
        check(rc, "ostp_init")

    def vec3(self, n: int = 1, single: bool = False) -> np.ndarray:
        dt = np.complex64 if single else np.complex128
        return np.zeros((n, 3, self.sizeh) if n > 1 else (3, self.sizeh), dtype=dt)

    def gauge(self, single: bool = False) -> np.ndarray:
        """su3_soa[8] as raw bytes (K plane included when par_temp)."""
        nb = 8 * (self.su3f_bytes if single else self.su3_bytes)
        return np.zeros(nb, dtype=np.uint8)

    def links(self, g: np.ndarray, single: bool = False) -> np.ndarray:
        """(8, 3, 3, sizeh) complex view of the 9 colour planes of each link array."""
        dt = np.dtype(np.complex64 if single else np.complex128)
        stride = self.su3f_bytes if single else self.su3_bytes
        base = g.view(np.uint8)
        return np.lib.stride_tricks.as_strided(
            base.view(dt) if stride % dt.itemsize == 0 else base.view(dt),
            shape=(8, 3, 3, self.sizeh),
            strides=(stride, 3 * self.sizeh * dt.itemsize, self.sizeh * dt.itemsize, dt.itemsize),
            writeable=True,
        )

    def K(self, g: np.ndarray) -> np.ndarray:
        assert self.par_temp
        base = g.view(np.uint8)
        return np.lib.stride_tricks.as_strided(
            base[144 * self.sizeh:].view(np.float64),
            shape=(8, self.sizeh),
            strides=(self.su3_bytes, 8),
            writeable=True,
        )

    def phases(self, single: bool = False) -> np.ndarray:
        return np.zeros((8, self.sizeh), dtype=np.float32 if single else np.float64)


class Present:
    """A host array mirrored on the device (the replacement of
    ``#pragma acc enter data copyin`` / ``update device`` / ``update self``)."""

    def __init__(self, a: np.ndarray, copyin: bool = True):
        assert a.flags.c_contiguous
        self.a = a
        fn = lib.ostp_acc_copyin if copyin else lib.ostp_acc_create
        check(fn(ptr(a), a.nbytes), "acc_copyin")

    @property
    def p(self) -> int:
        return ptr(self.a)

    def update_device(self) -> None:
        check(lib.ostp_acc_update_device(self.p, self.a.nbytes), "acc_update_device")

    def update_self(self) -> np.ndarray:
        check(lib.ostp_acc_update_self(self.p, self.a.nbytes), "acc_update_self")
        return self.a

    def delete(self) -> None:
        lib.ostp_acc_delete(self.p)


def ferm_param(mass: float, phases_ptr: int, phases_f_ptr: int | None = None) -> FermParam:
    fp = FermParam()
    fp.ferm_mass = mass
    fp.phases = phases_ptr
    fp.phases_f = phases_f_ptr or 0
    fp.ferm_charge = 0.0
    fp.degeneracy = 1
    fp.number_of_ps = 1
    return fp


def rational_approx(a0: float, a, b, *, num: int = -1, den: int = 4, lmin: float = 0.0, lmax: float = 1.0,
                    error: float = 0.0) -> RationalApprox:
    ra = RationalApprox()
    ra.exponent_num, ra.exponent_den = num, den
    ra.approx_order = len(a)
    ra.lambda_min, ra.lambda_max, ra.error = lmin, lmax, error
    ra.RA_a0 = a0
    for k, (ak, bk) in enumerate(zip(a, b)):
        ra.RA_a[k] = ak
        ra.RA_b[k] = bk
    return ra


def read_remez(path: str) -> RationalApprox:
    """Parse a .REMEZ file (format of rationalapprox_read_custom_nomefile,
    RationalApprox/rationalapprox.c:83-117)."""
    import re

    txt = pathlib.Path(path).read_text()
    num, den = map(int, re.search(r"\(x\)\^\((-?\d+)/(-?\d+)\)", txt).groups())
    order = int(re.search(r"Order:\s*(\d+)", txt).group(1))
    lmin = float(re.search(r"Lambda Min:\s*(\S+)", txt).group(1))
    lmax = float(re.search(r"Lambda Max:\s*(\S+)", txt).group(1))
    err = float(re.search(r"Error:\s*(\S+)", txt).group(1))
    a0 = float(re.search(r"RA_a0 = (\S+)", txt).group(1))
    pairs = re.findall(r"RA_a\[(\d+)\] = ([^,]+), RA_b\[\d+\] = (\S+)", txt)
    if len(pairs) != order:
        raise ValueError(f"{path}: expected {order} poles, found {len(pairs)}")
    a = [float(p[1]) for p in pairs]
    b = [float(p[2]) for p in pairs]
    ra = rational_approx(a0, a, b, num=num, den=den, lmin=lmin, lmax=lmax, error=err)
    return ra
