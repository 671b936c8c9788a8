"""Fermion measurements: the mu-derivative operators, the field-weighted
Dslash, idphase_dbz and the batched fermion_measures driver.

Bar:
  * dM_dmu_eo/oe[mu], d2M_dmu2_eo/oe[mu] (baryon_number_utilities.c) against
    the reference's own functions compiled from that file (oracle/_ref), with
    electromagnetic and imaginary-chemical-potential phases, to 1e-14 relative;
  * acc_Deo_wf / acc_Doe_wf (field_times_fermion_matrix.c:77-236, a file that
    needs generated headers and does not build here), pinned by what fixes the
    weighted hop: w = 1 gives acc_Deo / acc_Doe; a real 0/1 weight on a random
    half of the links gives the oracle's Dslash of the field with the other
    links zeroed; an imaginary weight i a is the derivative of the Dslash along
    the phases a (central difference of the oracle's Dslash, h = 1e-4); the
    weight enters real-linearly;
  * idphase_dbz (magnetic_susceptibility_utilities.c:13-35) against the same
    steps composed from the reference's compiled backfield.c routines, with a
    direction map that moves time;
  * ostp_fermion_measures (ferm_meas.c:282-440 for one flavour) against the
    oracle's eo_inversion (1e-12), the reference's compiled derivative
    operators and the central-difference weighted Dslash on the same Z2 noise,
    every one of the six measures, to 1e-6 relative; measures not asked for
    are NaN.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import openstaple_amd as ost  # noqa: E402
from openstaple_amd import lib  # noqa: E402
from conftest import ref_lib  # noqa: E402
from gpu_util import Dev, init_lattice, rel  # noqa: E402
from oracle import Oracle  # noqa: E402

BF = (0.0, 0.5, 0.0, 1.0, 0.0, -2.0)


def derivative(kind, eo, mu):
    name = ("dM_dmu_%s%d" if kind == 0 else "d2M_dmu2_%s%d") % ("eo" if eo == 0 else "oe", mu)
    return getattr(lib, name)


@pytest.fixture
def identity_map():
    yield
    lib.ostp_set_direction_map(0, 1, 2, 3)


@pytest.mark.parametrize("shape", [(4, 4, 4, 4), (4, 6, 8, 10)])
def test_derivatives_vs_reference(shape):
    ref = ref_lib(shape)
    if ref is None:
        pytest.skip("oracle/_ref not built for %s" % (shape,))
    o = Oracle(*shape)
    init_lattice(shape)
    u = o.gen_expH(0.4, 2)
    ph = o.phases(bf=BF, mu=0.3, charge=1.0)
    inp = o.gen_gauss(9, 0)
    du, dph, din = Dev(u), Dev(ph), Dev(inp)
    dout = Dev(nbytes=inp.nbytes)
    for kind in (0, 1):
        for eo in (0, 1):
            for mu in range(4):
                expect = o.vec3()
                ref.ref_dmu(kind, eo, mu, C.c_void_p(u.ctypes.data), C.c_void_p(expect.ctypes.data),
                            C.c_void_p(inp.ctypes.data), C.c_void_p(ph.ctypes.data))
                derivative(kind, eo, mu)(du.p, dout.p, din.p, dph.p)
                ost.check(0, "derivative")
                got = dout.get(expect)
                assert rel(got, expect) < 1e-14, (kind, eo, mu, rel(got, expect))
    # the tables index the same functions
    for mu in range(4):
        for tab, kind, eo in (("dM_dmu_eo", 0, 0), ("dM_dmu_oe", 0, 1), ("d2M_dmu2_eo", 1, 0),
                              ("d2M_dmu2_oe", 1, 1)):
            fp = (C.c_void_p * 4).in_dll(lib, tab)[mu]
            assert fp == C.cast(derivative(kind, eo, mu), C.c_void_p).value


def wf(eo, du, din, dph, wre, wim, like):
    out = Dev(nbytes=like.nbytes)
    fn = lib.acc_Deo_wf if eo == 0 else lib.acc_Doe_wf
    fn(du.p, out.p, din.p, dph.p, wre.p, wim.p)
    ost.check(0, "acc_D_wf")
    return out.get(like)


@pytest.mark.parametrize("shape", [(4, 6, 8, 10), (8, 8, 8, 8)])
def test_weighted_dslash(shape):
    o = Oracle(*shape)
    init_lattice(shape)
    rng = np.random.default_rng(5)
    u = o.gen_expH(0.3, 4)
    ph = o.phases(bf=BF, mu=0.2, charge=2.0 / 3)
    inp = o.gen_gauss(3, 1)
    du, dph, din = Dev(u), Dev(ph), Dev(inp)
    ones, zeros = np.ones((8, o.sizeh)), np.zeros((8, o.sizeh))
    mask = (rng.random((8, o.sizeh)) < 0.5).astype(np.float64)
    a = rng.normal(size=(8, o.sizeh))
    # the oracle's Dslash of the field with the links outside the mask zeroed
    planes = u.view(np.complex128).reshape(8, 9, o.sizeh)
    um = np.ascontiguousarray((planes * mask[:, None, :]).reshape(-1)).view(np.uint8)
    h = 1e-4
    for eo in (0, 1):
        D = o.deo if eo == 0 else o.doe
        # w = 1
        got = wf(eo, du, din, dph, Dev(ones), Dev(zeros), inp)
        assert rel(got, D(u, inp, ph)) < 1e-14
        # real 0/1 weights
        got_m = wf(eo, du, din, dph, Dev(mask), Dev(zeros), inp)
        assert rel(got_m, D(um, inp, ph)) < 1e-14
        # imaginary weights: d/dh D(phases + h a)
        got_a = wf(eo, du, din, dph, Dev(zeros), Dev(a), inp)
        fd = (D(u, inp, ph + h * a) - D(u, inp, ph - h * a)) / (2 * h)
        assert rel(got_a, fd) < 1e-7, rel(got_a, fd)
        # real-linearity in the weight
        got_c = wf(eo, du, din, dph, Dev(2.5 * mask), Dev(a), inp)
        assert rel(got_c, 2.5 * got_m + got_a) < 1e-14


def test_weighted_dslash_needs_both_parts():
    shape = (4, 4, 4, 4)
    o = Oracle(*shape)
    init_lattice(shape)
    u, ph, inp = o.gen_expH(0.3, 1), o.phases(), o.gen_gauss(1, 0)
    du, dph, din, dw = Dev(u), Dev(ph), Dev(inp), Dev(np.ones((8, o.sizeh)))
    out = Dev(nbytes=inp.nbytes)
    lib.acc_Deo_wf(du.p, out.p, din.p, dph.p, dw.p, None)
    assert b"go together" in lib.ostp_last_error()
    lib.ostp_clear_error()


@pytest.mark.parametrize("charge", [1.0, -1.0 / 3])
def test_idphase_dbz_vs_reference(identity_map, charge):
    shape = (4, 6, 8, 10)
    ref = ref_lib(shape)
    if ref is None:
        pytest.skip("oracle/_ref not built for %s" % (shape,))
    o = Oracle(*shape)
    init_lattice(shape)
    m = (1, 3, 0, 2)
    assert lib.ostp_set_direction_map(*m) == 0
    er, ei = np.full((8, o.sizeh), 7.0), np.zeros((8, o.sizeh))
    ref.ref_idphase_dbz(C.c_void_p(er.ctypes.data), C.c_void_p(ei.ctypes.data), (C.c_int * 4)(*m),
                        C.c_double(charge))
    dre, dim = Dev(np.full((8, o.sizeh), 3.0)), Dev(nbytes=ei.nbytes)
    fp = ost.FermParam()
    fp.ferm_charge = charge
    lib.idphase_dbz(dre.p, dim.p, C.byref(fp))
    ost.check(0, "idphase_dbz")
    assert np.all(dre.get(er) == 0.0) and np.all(er == 0.0)
    got = dim.get(ei)
    assert np.max(np.abs(ei)) > 0.1
    assert np.max(np.abs(got - ei)) <= 1e-13 * np.max(np.abs(ei))


def cpu_measures(o, ref, u, ph, mag_im, mass, deg, seed, nvec, n_chi, n_qn, tdir):
    """the six measures of ferm_meas.c:309-438 on the CPU, (nvec, 6) complex"""
    vol = float(np.prod(o.nd))
    factor = deg * 0.25 / vol
    h = 1e-4
    out = np.full((nvec, 6), np.nan + 1j * np.nan)

    def dmu(kind, ve, vo):
        ae, ao = o.vec3(), o.vec3()
        for eo, dst, src in ((0, ae, vo), (1, ao, ve)):
            ref.ref_dmu(kind, eo, tdir, C.c_void_p(u.ctypes.data), C.c_void_p(dst.ctypes.data),
                        C.c_void_p(src.ctypes.data), C.c_void_p(ph.ctypes.data))
        return ae, ao

    def inv(be, bo):
        st, _, xe, xo = o.eo_inversion(u, ph, mass, be, bo, 1e-12)
        return xe, xo

    for k in range(nvec):
        re_, ro_ = o.gen_z2(seed, 2 * k), o.gen_z2(seed, 2 * k + 1)

        def dot(ae, ao):
            return (np.vdot(re_, ae) + np.vdot(ro_, ao)) * factor

        xe, xo = inv(re_, ro_)
        out[k, 0] = dot(xe, xo)
        bne, bno = dmu(0, xe, xo)
        out[k, 1] = dot(bne, bno)
        if mag_im is not None:
            me = (o.deo(u, xo, ph + h * mag_im) - o.deo(u, xo, ph - h * mag_im)) / (2 * h)
            mo = (o.doe(u, xe, ph + h * mag_im) - o.doe(u, xe, ph - h * mag_im)) / (2 * h)
            out[k, 2] = -dot(me, mo)
        if k < n_chi:
            ye, yo = inv(xe, xo)
            out[k, 3] = -dot(ye, yo)
        if k < n_qn:
            ce, co = dmu(1, xe, xo)
            out[k, 4] = dot(ce, co)
            ye, yo = inv(bne, bno)
            de, do = dmu(0, ye, yo)
            out[k, 5] = -dot(de, do)
    return out


@pytest.mark.parametrize("half,with_mag,tmap", [(0, True, 3), (1, True, 0), (0, False, 3)])
def test_fermion_measures_vs_oracle(identity_map, half, with_mag, tmap):
    shape = (4, 6, 8, 10)
    ref = ref_lib(shape)
    if ref is None:
        pytest.skip("oracle/_ref not built for %s" % (shape,))
    o = Oracle(*shape)
    init_lattice(shape)
    m4 = (0, 1, 2, 3) if tmap == 3 else (1, 2, 3, 0)
    assert lib.ostp_set_direction_map(*m4) == 0
    mass, deg, seed, nvec, n_chi, n_qn = 0.1, 2, 41, 3, 2, 1
    u = o.gen_expH(0.35, 6)
    ph = o.phases(map4=m4, bf=BF, mu=0.25, charge=1.0)
    du, dph = Dev(u), Dev(ph)
    duf, dphf = Dev(o.to_float_gauge(u)), Dev(ph.astype(np.float32))
    fp = ost.ferm_param(mass, dph.p, dphf.p)
    fp.degeneracy = deg
    fp.ferm_charge = 1.0
    mag_im = None
    if with_mag:
        mre, mim = Dev(nbytes=ph.nbytes), Dev(nbytes=ph.nbytes)
        lib.idphase_dbz(mre.p, mim.p, C.byref(fp))
        ost.check(0, "idphase_dbz")
        mag_im = mim.get(ph)
        fp.mag_re, fp.mag_im = mre.p, mim.p
    res = (C.c_double * (12 * nvec))()
    cg = C.c_int(0)
    conv = lib.ostp_fermion_measures(du.p, duf.p, C.byref(fp), nvec, n_chi, n_qn, seed, 1e-10, 10000, half, res,
                                     C.byref(cg))
    ost.check(0, "ostp_fermion_measures")
    assert conv == nvec and cg.value > 0
    got = np.array(res[0::2]).reshape(nvec, 6) + 1j * np.array(res[1::2]).reshape(nvec, 6)
    expect = cpu_measures(o, ref, u, ph, mag_im, mass, deg, seed, nvec, n_chi, n_qn, m4[3])
    scale = np.max(np.abs(expect[:, 0]))
    for k in range(nvec):
        for q in range(6):
            e, g = expect[k, q], got[k, q]
            if np.isnan(e):
                assert np.isnan(g.real) and np.isnan(g.imag), (k, q, g)
            else:
                assert abs(g - e) <= 1e-6 * max(abs(e), 1e-3 * scale), (k, q, g, e)


def test_fermion_measures_rejects_bad_arguments():
    shape = (4, 4, 4, 4)
    o = Oracle(*shape)
    init_lattice(shape)
    u, ph = o.gen_expH(0.3, 1), o.phases()
    du, dph = Dev(u), Dev(ph)
    fp = ost.ferm_param(0.1, dph.p, 0)
    res = (C.c_double * 12)()
    assert lib.ostp_fermion_measures(du.p, None, C.byref(fp), 0, 0, 0, 1, 1e-10, 100, 0, res, None) == 0
    assert b"invalid arguments" in lib.ostp_last_error()
    lib.ostp_clear_error()
