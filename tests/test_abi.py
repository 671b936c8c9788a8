"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import pathlib
import re
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADERS = sorted((ROOT / "include").glob("*.h"))
LIB = ROOT / "openstaple_amd" / "libopenstaple_amd.so"


def declared_functions():
    names = set()
    for h in HEADERS:
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"^typedef[^;]*;", "", txt, flags=re.M | re.S)  # type and callback typedefs
        txt = re.sub(r'^extern\s+(?!"C")[^;]*;', "", txt, flags=re.M | re.S)  # global variables
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof", "defined"):
                continue
            names.add(name)
    return names


@pytest.fixture(scope="module")
def lib():
    if not LIB.exists():
        pytest.skip("library not built")
    return ctypes.CDLL(str(LIB))


def test_headers_declare_the_reference_entry_points():
    names = declared_functions()
    for ref in ["acc_Deo", "acc_Doe", "fermion_matrix_multiplication", "fermion_matrix_multiplication_shifted",
                "ker_invert_openacc", "inverter_mixed_precision", "multishift_invert",
                "recombine_shifted_vec3_to_vec3", "inverter_wrapper", "inverter_multishift_wrapper",
                "calc_u1_phases", "l2norm2_global", "real_scal_prod_global"]:
        assert ref in names, ref


def test_library_exports_every_declared_function(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = sorted(n for n in declared_functions() if n not in exported)
    assert not missing, missing


def test_reference_globals_are_exported(lib):
    for g in ("verbosity_lv", "inverter_tricks", "multishift_invert_iterations", "deltas_Omelyan", "md_parameters",
              "act_params", "nMdInversionPerformed", "conf_acc_bkp", "aux_conf_acc", "auxbis_conf_acc",
              "gstout_conf_acc_arr", "gconf_as_fermionmatrix", "momenta", "ipdot_acc", "ferm_chi_acc",
              "ferm_phi_acc", "ferm_shiftmulti_acc", "kloc_r", "kloc_h", "kloc_s", "kloc_p", "k_p_shiftferm",
              "local_sums", "d_local_sums", "fermions_parameters", "alloc_info", "input_file_str", "dM_dmu_eo",
              "dM_dmu_oe", "d2M_dmu2_eo", "d2M_dmu2_oe"):
        ctypes.c_int.in_dll(lib, g)
    # the derivative tables point at the exported functions
    for name in ("dM_dmu_eo", "dM_dmu_oe", "d2M_dmu2_eo", "d2M_dmu2_oe"):
        table = (ctypes.c_void_p * 4).in_dll(lib, name)
        for k in range(4):
            assert table[k] == ctypes.cast(getattr(lib, "%s%d" % (name, k)), ctypes.c_void_p).value


def test_python_binding_covers_the_exports():
    import openstaple_amd as ost

    missing = sorted(n for n in declared_functions() if n not in ost.EXPORTED)
    assert not missing, missing


def test_c_host_compiles_and_links_against_the_header(tmp_path):
    """A plain C host (like the reference's) includes the header and links."""
    src = tmp_path / "host.c"
    src.write_text(
        '#include "openstaple_amd_types.h"\n'
        "int main(void){ ferm_param fp; inverter_package ip; RationalApprox ra;\n"
        "  (void)fp; (void)ip; (void)ra;\n"
        "  return (int)(sizeof(RationalApprox) != 456) + (int)(sizeof(inverter_package) != 112); }\n"
    )
    exe = tmp_path / "host"
    subprocess.run(["gcc", "-std=gnu99", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
                    f"-L{LIB.parent}", "-lopenstaple_amd", f"-Wl,-rpath,{LIB.parent}"], check=True)
