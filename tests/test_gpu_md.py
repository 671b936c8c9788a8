"""Molecular-dynamics pieces on the GPU: momenta, mom_sum_mult,
mom_exp_times_conf_soloopenacc, calc_momenta_action, invert_momenta and the
gauge sub-integrator multistep_2MN_gauge.

Bar:
  * against the reference's own su3_utilities.c / su3_measurements.c /
    md_parameters.c compiled in oracle/_ref: mom_sum_mult bit-exact,
    the link update and the momenta action 1e-13, the whole gauge
    sub-integrator (composed from the compiled routines in the order of
    md_integrator.c:347-376) 1e-11;
  * the full 2MN trajectory with the stouted fermion force
    (multistep_2MN_SOLOOPENACC) is reversible to the solver tolerance and
    conserves H = P^2/2 + S_gauge + <chi, r(M^+M) chi> to second order in the
    outer step (ratio 4 within 20% when the step halves), with r the same
    rational function as the force;
  * independently of the reference: the integrator is reversible (forward,
    P -> -P, forward returns to the start, 1e-11) and its energy violation
    falls as the square of the step (ratio 4 within 12% when the step halves),
    which fixes the relative normalisation of the force, the momenta action
    and the step sizes; the generated momenta have the second moments of
    the reference's construction (<rc00^2> = 2/3, <rc00 rc11> = -1/3,
    <|c_ij|^2> = 1, Tr(P^2)/2 = 4 per link) within 5 standard errors.
"""
import ctypes as C
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import openstaple_amd as ost  # noqa: E402
from openstaple_amd import lib  # noqa: E402
from conftest import ref_lib  # noqa: E402
from gpu_util import Dev, init_lattice, rel  # noqa: E402
from oracle import Oracle  # noqa: E402

C0_TLSM, C1_TLSM = 5.0 * 0.33333333333333333333333, -0.25 * 0.33333333333333333333333
GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def vp(a):
    return C.c_void_p(a.ctypes.data)


def setup(shape, seed, tlsm=True):
    o = Oracle(*shape, par_temp=False)
    u = o.gen_expH(0.4, seed)
    init_lattice(shape)
    lib.ostp_set_stout(0.15, 2, 0.0, 0, 0, int(tlsm))
    return o, u


def momenta(sizeh, seed, stream=0):
    dm = Dev(nbytes=8 * 64 * sizeh)
    lib.ostp_generate_Momenta_gauss(dm.p, seed, stream)
    ost.check(0, "ostp_generate_Momenta_gauss")
    return dm.get(np.zeros(8 * 8 * sizeh))


def set_md(beta, no_md, gauge_scale, t):
    ost.ActionParam.in_dll(lib, "act_params").beta = beta
    p = ost.MdParam(no_md=no_md, gauge_scale=gauge_scale, t=t)
    mp = ost.MdParam.in_dll(lib, "md_parameters")
    C.memmove(C.addressof(mp), C.addressof(p), C.sizeof(p))
    lib.initialize_md_global_variables(p)
    return np.ctypeslib.as_array((C.c_double * 7).in_dll(lib, "deltas_Omelyan")).copy()


@pytest.mark.parametrize("shape", [(4, 4, 4, 4), (4, 6, 8, 10)])
def test_momentum_pieces_vs_reference_build(shape):
    ref = ref_lib(shape)
    if ref is None:
        pytest.skip("oracle/_ref not built")
    ref.ref_momenta_action.restype = C.c_double
    o, u = setup(shape, 7)
    sizeh = o.sizeh
    mom = momenta(sizeh, 11)
    rng = np.random.default_rng(3)
    ipd = rng.normal(size=8 * 8 * sizeh)
    factor = np.array([0.0, 0.0, -0.0371, 0.0, 0.0, 0.0, 0.0])

    dm, di = Dev(mom), Dev(ipd)
    lib.mom_sum_mult(dm.p, di.p, factor.ctypes.data, 2)
    ost.check(0, "mom_sum_mult")
    got = dm.get(mom)
    want = mom.copy()
    ref.ref_mom_sum_mult(vp(want), vp(ipd), C.c_double(factor[2]))
    assert np.array_equal(got, want)

    du = Dev(u)
    delta = np.array([0.0, 0.0, 0.0, 0.0, 0.113, 0.0, 0.0])
    lib.mom_exp_times_conf_soloopenacc(du.p, dm.p, delta.ctypes.data, 4)
    ost.check(0, "mom_exp_times_conf_soloopenacc")
    ug = du.get(u)
    ur = u.copy()
    ref.ref_mom_exp(vp(ur), vp(want), C.c_double(delta[4]))
    assert rel(ug.view(np.float64), ur.view(np.float64)) < 1e-13
    assert not np.array_equal(ur, u)

    for mu in range(8):
        a = lib.calc_momenta_action(dm.p, None, mu)
        b = ref.ref_momenta_action(vp(want), mu)
        assert abs(a - b) <= 1e-13 * abs(b), (mu, a, b)

    lib.invert_momenta(dm.p)
    assert np.array_equal(dm.get(mom), -got)


@pytest.mark.parametrize("shape,no_md,gs", [((4, 4, 4, 4), 2, 2), ((4, 6, 8, 10), 1, 3)])
def test_gauge_integrator_vs_reference_build(shape, no_md, gs):
    ref = ref_lib(shape)
    if ref is None:
        pytest.skip("oracle/_ref not built")
    o, u = setup(shape, 5)
    sizeh = o.sizeh
    beta, t = 3.8, 0.3
    deltas = set_md(beta, no_md, gs, t)
    rd = np.zeros(7)
    ref.ref_md_deltas(C.c_double(beta), no_md, gs, C.c_double(t), vp(rd))
    assert np.array_equal(deltas, rd)
    mom = momenta(sizeh, 2)
    du, dm = Dev(u), Dev(mom)
    ds = Dev(nbytes=u.nbytes)
    di = Dev(nbytes=8 * 64 * sizeh)
    lib.multistep_2MN_gauge(du.p, ds.p, di.p, dm.p)
    ost.check(0, "multistep_2MN_gauge")
    ur, mr = u.copy(), mom.copy()
    sr = np.zeros_like(u)
    ir = np.zeros(8 * 8 * sizeh)
    ref.ref_md_gauge(vp(ur), vp(sr), vp(ir), vp(mr), C.c_double(beta), no_md, gs, C.c_double(t))
    assert rel(du.get(u).view(np.float64), ur.view(np.float64)) < 1e-11
    assert rel(dm.get(mom), mr) < 1e-11
    assert rel(di.get(ir), ir) < 1e-11
    assert rel(ds.get(sr).view(np.float64), sr.view(np.float64)) < 1e-11


def hamiltonian(u_dev, mom_dev, beta, tlsm=True):
    p = lib.calc_plaquette_soloopenacc(u_dev.p, None, None)
    s = (C0_TLSM if tlsm else 1.0) * p
    if tlsm:
        s += C1_TLSM * lib.calc_rettangolo_soloopenacc(u_dev.p, None, None)
    kin = sum(lib.calc_momenta_action(mom_dev.p, None, mu) for mu in range(8))
    return kin - beta / 3.0 * s


@pytest.mark.parametrize("tlsm", [1, 0])
def test_gauge_integrator_reversible_and_second_order(tlsm):
    shape = (4, 4, 4, 8)
    o, u = setup(shape, 9, tlsm=bool(tlsm))
    sizeh = o.sizeh
    beta, t = 5.0, 0.5
    mom = momenta(sizeh, 4)

    # reversibility
    set_md(beta, 1, 4, t)
    du, dm = Dev(u), Dev(mom)
    ds, di = Dev(nbytes=u.nbytes), Dev(nbytes=8 * 64 * sizeh)
    lib.multistep_2MN_gauge(du.p, ds.p, di.p, dm.p)
    moved = du.get(u)
    assert rel(moved.view(np.float64), u.view(np.float64)) > 1e-3
    lib.invert_momenta(dm.p)
    lib.multistep_2MN_gauge(du.p, ds.p, di.p, dm.p)
    ost.check(0, "multistep_2MN_gauge")
    assert rel(du.get(u).view(np.float64), u.view(np.float64)) < 1e-11
    assert rel(dm.get(mom), -mom) < 1e-11

    # energy violation ~ step^2 (time t/2 in gauge_scale 2MN steps)
    dh = []
    for gs in (8, 16):
        set_md(beta, 1, gs, t)
        du, dm = Dev(u), Dev(mom)
        h0 = hamiltonian(du, dm, beta, bool(tlsm))
        lib.multistep_2MN_gauge(du.p, ds.p, di.p, dm.p)
        dh.append(hamiltonian(du, dm, beta, bool(tlsm)) - h0)
    ratio = dh[0] / dh[1]
    assert abs(ratio - 4.0) < 0.48, (dh, ratio)


def test_momenta_statistics_and_streams():
    shape = (8, 8, 8, 16)
    init_lattice(shape)
    sizeh = 8 * 8 * 8 * 16 // 2
    m = momenta(sizeh, 21).reshape(8, 8 * sizeh)
    again = momenta(sizeh, 21).reshape(8, 8 * sizeh)
    assert np.array_equal(m, again)
    other = momenta(sizeh, 21, stream=1).reshape(8, 8 * sizeh)
    assert not np.any(m == other)
    # thmat_soa: c01, c02, c12 complex planes, then rc00, rc11
    c = np.ascontiguousarray(m[:, :6 * sizeh]).view(np.complex128).reshape(8, 3, sizeh)
    rc00, rc11 = m[:, 6 * sizeh:7 * sizeh], m[:, 7 * sizeh:]
    n = rc00.size

    def near(x, mean, var):
        se = np.sqrt(var / n)
        assert abs(x.mean() - mean) < 5 * se, (x.mean(), mean, se)

    near(rc00 ** 2, 2 / 3, 2 * (2 / 3) ** 2)
    near(rc11 ** 2, 2 / 3, 2 * (2 / 3) ** 2)
    near(rc00 * rc11, -1 / 3, (2 / 3) ** 2 + (1 / 3) ** 2)
    for k in range(3):
        near(np.abs(c[:, k]) ** 2, 1.0, 2 * 0.25 * 2)
        near(c[:, k].real * c[:, k].imag, 0.0, 0.25)
    half_tr = rc00 ** 2 + rc11 ** 2 + rc00 * rc11 + (np.abs(c) ** 2).sum(axis=1)
    near(half_tr, 4.0, 8 * 2 * 0.25)
    dm = Dev(m.ravel())
    tot = sum(lib.calc_momenta_action(dm.p, None, mu) for mu in range(8))
    assert abs(tot - half_tr.sum()) <= 1e-12 * tot


def approx_md9():
    """x^(-1/4) of order 9 on [32 lambda_min, 32], from a committed .REMEZ file"""
    ra = ost.read_remez(str(GOLDEN / "approx_-1_over_4_order_9_mloglm_6.4.REMEZ"))
    lmax = 32.0
    eps = lmax ** (ra.exponent_num / ra.exponent_den)
    ra.RA_a0 *= eps
    for k in range(ra.approx_order):
        ra.RA_a[k] *= lmax * eps
        ra.RA_b[k] *= lmax
    return ra


class FermionMD:
    """One staggered flavour (stout 2 x 0.15, tree-level Symanzik gauge action)
    with nps pseudofermions chi, the buffers multistep_2MN_SOLOOPENACC takes,
    and the fermion action <chi, r(M^+M) chi> through the library's own
    multishift solver and recombination."""

    def __init__(self, shape, seed, mass=0.1, nps=2, res=1e-12, bf=(0, 0, 0, 0, 0, 0), mu=0.0, charge=0.0):
        o = Oracle(*shape)
        init_lattice(shape)
        lib.ostp_set_stout(0.15, 2, 0.0, 0, 0, 1)
        lib.ostp_set_inverter_tricks(0, 0, 0.1, 10000)
        self.o, self.nps, self.res = o, nps, res
        self.u = o.gen_expH(0.3, seed)
        self.lat = lat = ost.Lattice(*shape)
        self.ph = Dev(o.phases(bf=bf, mu=mu, charge=charge))
        self.ra = approx_md9()
        order = self.ra.approx_order
        self.chi = Dev(np.stack([o.gen_gauss(90 + i) for i in range(nps)]))
        gb = self.u.nbytes
        self.arr = Dev(nbytes=2 * gb)
        self.gl3, self.taux = Dev(nbytes=gb), Dev(nbytes=gb)
        self.ipd = Dev(nbytes=8 * 64 * o.sizeh)
        self.xs = Dev(nbytes=order * lat.vec3_bytes)
        self.sf = Dev(nbytes=order * lat.vec3_bytes)
        self.d64 = [Dev(nbytes=lat.vec3_bytes) for _ in range(4)]
        self.out = Dev(nbytes=lat.vec3_bytes)
        self.fps = (ost.FermParam * 1)()
        self.fps[0].ferm_mass = mass
        self.fps[0].degeneracy = 1
        self.fps[0].phases = self.ph.p
        self.fps[0].number_of_ps = nps
        self.fps[0].index_of_the_first_ps = 0
        self.fps[0].approx_md = self.ra

    def ip(self, conf_p):
        ipk = ost.InverterPackage()
        ipk.u = conf_p
        ipk.ferm_shift_temp, ipk.nshifts = self.sf.p, self.ra.approx_order
        ipk.loc_r, ipk.loc_h, ipk.loc_s, ipk.loc_p = (d.p for d in self.d64)
        return ipk

    def trajectory(self, du, dm):
        lib.multistep_2MN_SOLOOPENACC(self.ipd.p, du.p, self.arr.p, self.gl3.p, self.taux.p, self.fps, 1,
                                      self.chi.p, self.xs.p, self.ip(du.p), dm.p, None, self.res, 10000)
        ost.check(0, "multistep_2MN_SOLOOPENACC")

    def action(self, du):
        lib.stout_wrapper(du.p, self.arr.p, 0)
        v = self.arr.p + self.u.nbytes
        tot = 0.0
        for i in range(self.nps):
            chi_i = self.chi.p + i * self.lat.vec3_bytes
            lib.inverter_multishift_wrapper(self.ip(v), self.fps, C.addressof(self.ra), self.xs.p, chi_i,
                                            self.res, 10000, 1)
            lib.recombine_shifted_vec3_to_vec3(self.xs.p, chi_i, self.out.p, C.addressof(self.ra))
            tot += lib.real_scal_prod_global(chi_i, self.out.p)
        ost.check(0, "fermion action")
        return tot


# no background field (phases +-1), and an e.m. field with an imaginary
# chemical potential (complex phases in the Dslash and in the force)
@pytest.mark.parametrize("field", [dict(), dict(bf=(0.7, -1.3, 0.4, 1.0, 3.0, -2.0), mu=0.25, charge=2.0 / 3.0)],
                         ids=["no_field", "em_field_imag_mu"])
def test_full_md_trajectory_reversible_and_second_order(field):
    shape = (4, 4, 4, 8)
    F = FermionMD(shape, 13, **field)
    sizeh = F.o.sizeh
    beta, t = 5.0, 0.5
    mom = momenta(sizeh, 6)
    u = F.u

    set_md(beta, 2, 2, t)
    du, dm = Dev(u), Dev(mom)
    F.trajectory(du, dm)
    assert rel(du.get(u).view(np.float64), u.view(np.float64)) > 1e-3
    lib.invert_momenta(dm.p)
    F.trajectory(du, dm)
    assert rel(du.get(u).view(np.float64), u.view(np.float64)) < 1e-9
    assert rel(dm.get(mom), -mom) < 1e-9

    def H(du, dm):
        return hamiltonian(du, dm, beta, True) + F.action(du)

    dh = []
    for no_md in (3, 6):
        set_md(beta, no_md, 2, t)
        du, dm = Dev(u), Dev(mom)
        h0 = H(du, dm)
        F.trajectory(du, dm)
        dh.append(H(du, dm) - h0)
    ratio = dh[0] / dh[1]
    assert abs(ratio - 4.0) < 0.8, (dh, ratio)
