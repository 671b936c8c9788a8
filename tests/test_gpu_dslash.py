"""GPU parity of the staggered Dslash, M^dag M, phase and neighbour tables,
through the C ABI, against the CPU oracle on the same inputs.

Bar: indexing and phase tables bit-exact; FP64 Dslash within 1e-13 relative
(norm-wise and element-wise) of the oracle; FP32 Dslash within 2e-6 of the FP32
oracle.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import openstaple_amd as ost  # noqa: E402
from openstaple_amd import lib  # noqa: E402
from gpu_util import Dev, init_lattice, rel  # noqa: E402
from oracle import Oracle  # noqa: E402

TOL64 = 1e-13
TOL32 = 2e-6

SHAPES = [(4, 4, 4, 4), (8, 8, 8, 8), (4, 6, 8, 10), (16, 4, 6, 2)]


def fields(o, seed, eps=0.5):
    u = o.gen_cold(eps, seed)
    psi = o.gen_gauss(seed + 1)
    return u, psi


def check64(got, ref):
    assert rel(got, ref) < TOL64
    assert np.max(np.abs(got - ref)) <= TOL64 * np.max(np.abs(ref))


@pytest.mark.parametrize("shape", SHAPES)
def test_nn_tables_bitexact(shape):
    o = Oracle(*shape)
    init_lattice(shape)
    nnp, nnm = o.nn_tables()
    gp = np.zeros_like(nnp)
    gm = np.zeros_like(nnm)
    ost.check(lib.ostp_compute_nnp_and_nnm(gp.ctypes.data, gm.ctypes.data), "nn")
    np.testing.assert_array_equal(gp, nnp)
    np.testing.assert_array_equal(gm, nnm)


PHASE_CASES = [
    ((0, 1, 2, 3), (0, 0, 0, 0, 0, 0), 0.0, 0.0),
    ((0, 1, 2, 3), (0.7, -1.3, 2.0, 1.0, 3.0, -2.0), 0.2, 2.0 / 3.0),
    ((3, 2, 0, 1), (1.0, 2.0, 3.0, 4.0, 5.0, 6.0), 0.35, -1.0 / 3.0),
]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("case", PHASE_CASES)
def test_phases_bitexact(shape, case):
    maps, bf, mu, charge = case
    o = Oracle(*shape)
    init_lattice(shape)
    ost.check(lib.ostp_set_direction_map(*maps), "map")
    ref = o.phases(maps, bf, mu, charge)
    host = np.zeros_like(ref)
    pr = ost.Present(host, copyin=False)
    lib.calc_u1_phases(pr.p, ost.BfParam(*bf), mu, charge)
    ost.check(0, "calc_u1_phases")
    np.testing.assert_array_equal(host.view(np.uint64), ref.view(np.uint64))
    dev = pr.update_self()
    np.testing.assert_array_equal(dev.view(np.uint64), ref.view(np.uint64))
    pr.delete()
    ost.check(lib.ostp_set_direction_map(0, 1, 2, 3), "map")


@pytest.mark.parametrize("shape", SHAPES)
def test_dslash_fp64_device_pointers(shape):
    o = Oracle(*shape)
    init_lattice(shape)
    u, psi = fields(o, 17)
    ph = o.phases(bf=(0.0, 0.5, 0.0, 1.0, 0.0, -2.0), mu=0.3, charge=1.0)
    du, dp, dpsi, dout = Dev(u), Dev(ph), Dev(psi), Dev(nbytes=psi.nbytes)
    lib.acc_Deo(du.p, dout.p, dpsi.p, dp.p)
    ost.check(0, "acc_Deo")
    check64(dout.get(psi), o.deo(u, psi, ph))
    lib.acc_Doe(du.p, dout.p, dpsi.p, dp.p)
    ost.check(0, "acc_Doe")
    check64(dout.get(psi), o.doe(u, psi, ph))


def test_dslash_fp64_present_host_pointers():
    """The drop-in path: host arrays registered like `acc enter data copyin`."""
    shape = (8, 8, 8, 8)
    o = Oracle(*shape)
    init_lattice(shape)
    u, psi = fields(o, 23)
    ph = o.phases()
    out = np.zeros_like(psi)
    pu, pp, pi, po = ost.Present(u), ost.Present(ph), ost.Present(psi), ost.Present(out, copyin=False)
    lib.acc_Doe(pu.p, po.p, pi.p, pp.p)
    ost.check(0, "acc_Doe")
    assert not out.any(), "host copy must not change before update self"
    po.update_self()
    check64(out, o.doe(u, psi, ph))
    for x in (pu, pp, pi, po):
        x.delete()


@pytest.mark.parametrize("shape", [(8, 8, 8, 8), (4, 6, 8, 10)])
def test_mdagm_shifted(shape):
    o = Oracle(*shape)
    init_lattice(shape)
    u, psi = fields(o, 31, eps=0.8)
    ph = o.phases()
    du, dp, dpsi = Dev(u), Dev(ph), Dev(psi)
    dout, dtmp = Dev(nbytes=psi.nbytes), Dev(nbytes=psi.nbytes)
    fp = ost.ferm_param(0.05, dp.p)
    lib.fermion_matrix_multiplication_shifted(du.p, dout.p, dpsi.p, dtmp.p, C.byref(fp), 0.013)
    ost.check(0, "mdagm")
    check64(dout.get(psi), o.mdagm(u, psi, ph, 0.05, 0.013))
    check64(dtmp.get(psi), o.doe(u, psi, ph))
    lib.fermion_matrix_multiplication(du.p, dout.p, dpsi.p, dtmp.p, C.byref(fp))
    check64(dout.get(psi), o.mdagm(u, psi, ph, 0.05, 0.0))


@pytest.mark.parametrize("shape", [(8, 8, 8, 8), (4, 6, 8, 10)])
def test_dslash_fp32(shape):
    o = Oracle(*shape)
    init_lattice(shape)
    u, psi = fields(o, 41)
    ph = o.phases()
    uf = o.to_float_gauge(u)
    phf = ph.astype(np.float32)
    psif = psi.astype(np.complex64)
    du, dp, dpsi, dout = Dev(uf), Dev(phf), Dev(psif), Dev(nbytes=psif.nbytes)
    lib.acc_Deo_f(du.p, dout.p, dpsi.p, dp.p)
    ost.check(0, "acc_Deo_f")
    assert rel(dout.get(psif), o.deo_f(uf, psif, phf)) < TOL32
    lib.acc_Doe_f(du.p, dout.p, dpsi.p, dp.p)
    assert rel(dout.get(psif), o.doe_f(uf, psif, phf)) < TOL32
    # and against FP64
    assert rel(dout.get(psif).astype(np.complex128), o.doe(u, psi, ph)) < 1e-6


def test_dslash_full_size_32():
    """32^4 (the benchmark lattice): GPU Deo/Doe vs the oracle on the same fields."""
    shape = (32, 32, 32, 32)
    o = Oracle(*shape)
    init_lattice(shape)
    u, psi = fields(o, 2, eps=0.5)
    ph = o.phases()
    du, dp, dpsi, dout = Dev(u), Dev(ph), Dev(psi), Dev(nbytes=psi.nbytes)
    lib.acc_Doe(du.p, dout.p, dpsi.p, dp.p)
    ost.check(0, "acc_Doe")
    check64(dout.get(psi), o.doe(u, psi, ph))
    lib.acc_Deo(du.p, dout.p, dpsi.p, dp.p)
    check64(dout.get(psi), o.deo(u, psi, ph))


def test_bad_pointer_is_an_error_not_a_fault():
    shape = (4, 4, 4, 4)
    o = Oracle(*shape)
    init_lattice(shape)
    u, psi = fields(o, 1)
    ph = o.phases()
    du, dp, d = Dev(u), Dev(ph), Dev(psi)
    stray = np.zeros_like(psi)  # host memory never registered
    lib.acc_Deo(du.p, d.p, stray.ctypes.data, dp.p)
    msg = lib.ostp_last_error()
    lib.ostp_clear_error()
    assert msg and b"neither present" in msg
    # a registered region too small for the field is rejected as well
    small = np.zeros(16, dtype=np.complex128)
    ps = ost.Present(small)
    lib.acc_Deo(du.p, d.p, ps.p, dp.p)
    msg = lib.ostp_last_error()
    lib.ostp_clear_error()
    assert msg and b"overruns" in msg
    ps.delete()
    # a device buffer too small for a gauge field
    lib.acc_Deo(d.p, d.p, d.p, dp.p)
    msg = lib.ostp_last_error()
    lib.ostp_clear_error()
    assert msg and b"overruns its allocation" in msg
