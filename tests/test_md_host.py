"""Host-side MD set-up: initialize_md_global_variables must give the
reference's deltas_Omelyan bit for bit (md_parameters.c:11-52, compiled in
oracle/_ref) for several step counts, gauge sub-step counts and couplings, and
the float copies must be their rounding.  No GPU is used."""
import ctypes as C

import numpy as np
import pytest

import openstaple_amd as ost
from openstaple_amd import lib
from conftest import ref_lib


@pytest.mark.parametrize("beta,no_md,gs,t", [(3.8, 10, 1, 1.0), (5.3, 7, 3, 0.85), (1.0, 1, 16, 0.5),
                                             (6.1, 13, 5, 2.0)])
def test_md_deltas_bit_exact(beta, no_md, gs, t):
    ref = ref_lib((4, 4, 4, 4))
    if ref is None:
        pytest.skip("oracle/_ref not built")
    ost.ActionParam.in_dll(lib, "act_params").beta = beta
    lib.initialize_md_global_variables(ost.MdParam(no_md=no_md, gauge_scale=gs, t=t))
    got = np.ctypeslib.as_array((C.c_double * 7).in_dll(lib, "deltas_Omelyan")).copy()
    got_f = np.ctypeslib.as_array((C.c_float * 7).in_dll(lib, "deltas_Omelyan_f")).copy()
    want = np.zeros(7)
    ref.ref_md_deltas(C.c_double(beta), no_md, gs, C.c_double(t), C.c_void_p(want.ctypes.data))
    assert np.array_equal(got, want)
    assert np.array_equal(got_f, want.astype(np.float32))
    # the Omelyan weights: lambda + (1 - 2 lambda) + lambda = 1 over a step
    lam = 0.1931833275037836
    assert got[0] / got[1] == pytest.approx(lam / (1 - 2 * lam), rel=1e-15)
    assert got[4] == pytest.approx(t / (4 * no_md * gs), rel=1e-15)


def test_md_param_layouts():
    assert C.sizeof(ost.MdParam) == 64 and ost.MdParam.extrapolateInvsForce.offset == 56
    assert C.sizeof(ost.ActionParam) == 80 and ost.ActionParam.topo_file_path.offset == 48
