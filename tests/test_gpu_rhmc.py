"""The RHMC update UPDATE_SOLOACC_UNOSTEP_VERSATILE on the GPU.

One staggered flavour (m = 0.1, two pseudofermions, stout 2 x 0.15, tree-level
Symanzik gauge action at beta = 5) on 4x4x4x8, mother approximations from the
committed .REMEZ files (x^(1/8) for the heat bath, x^(-1/4) of order 19 for
the MD and the final action).

Bar:
  * the initial actions the update reports are the gauge action
    beta/3 (c0 sum Re tr P + c1 sum Re tr R), Tr(P^2)/2 of the momenta and
    |phi|^2 of the pseudofermions, recomputed from the same generator streams,
    1e-12;
  * a trajectory of length zero changes the action by no more than the error
    of the rational approximations allows: |dS| <= 5 |phi|^2 (2 e_fi + e_li);
  * dS falls as the square of the MD step (ratio 4 within 25% when no_md
    doubles, same momenta and pseudofermions);
  * the Metropolis test: a uniform of 0 accepts and keeps the evolved links
    (bit-identical to the same trajectory run without the test), a uniform
    next to 1 rejects a trajectory with dS > 0 and restores the links bit for
    bit; the returned acceptance count follows.
"""
import ctypes as C
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import openstaple_amd as ost  # noqa: E402
from openstaple_amd import lib  # noqa: E402
from gpu_util import Dev, init_lattice, rel  # noqa: E402
from oracle import Oracle  # noqa: E402

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
C0_TLSM, C1_TLSM = 5.0 * 0.33333333333333333333333, -0.25 * 0.33333333333333333333333
BETA, MASS, NPS = 5.0, 0.1, 2
FI = "approx_1_over_8_order_19_mloglm_6.4.REMEZ"
LI = "approx_-1_over_4_order_19_mloglm_6.4.REMEZ"


def set_ptr(name, value):
    C.c_void_p.in_dll(lib, name).value = value


class Rhmc:
    def __init__(self, shape=(4, 4, 4, 8), seed=17):
        o = Oracle(*shape)
        self.o = o
        self.lat = lat = init_lattice(shape)
        lib.ostp_set_stout(0.15, 2, 0.0, 0, 0, 1)
        lib.ostp_set_inverter_tricks(0, 0, 0.1, 10000)
        ost.ActionParam.in_dll(lib, "act_params").beta = BETA
        self.u0 = o.gen_expH(0.3, seed)
        gb = self.u0.nbytes
        self.conf = Dev(self.u0)
        self.keep = [Dev(nbytes=n) for n in (gb, gb, gb, 2 * gb, 8 * 64 * o.sizeh, 8 * 64 * o.sizeh)]
        for name, d in zip(("conf_acc_bkp", "aux_conf_acc", "auxbis_conf_acc", "gstout_conf_acc_arr", "momenta",
                            "ipdot_acc"), self.keep):
            set_ptr(name, d.p)
        self.mom = self.keep[4]
        vb = lat.vec3_bytes
        self.phi, self.chi = Dev(nbytes=NPS * vb), Dev(nbytes=NPS * vb)
        self.shifts, self.kshift = Dev(nbytes=19 * vb), Dev(nbytes=19 * vb)
        self.kloc = [Dev(nbytes=vb) for _ in range(4)]
        set_ptr("ferm_phi_acc", self.phi.p)
        set_ptr("ferm_chi_acc", self.chi.p)
        set_ptr("ferm_shiftmulti_acc", self.shifts.p)
        set_ptr("k_p_shiftferm", self.kshift.p)
        for name, d in zip(("kloc_r", "kloc_h", "kloc_s", "kloc_p"), self.kloc):
            set_ptr(name, d.p)
        self.ph = Dev(o.phases())
        self.fps = (ost.FermParam * 1)()
        fp = self.fps[0]
        fp.ferm_mass, fp.degeneracy, fp.number_of_ps, fp.index_of_the_first_ps = MASS, 1, NPS, 0
        fp.phases = self.ph.p
        fp.approx_fi_mother = ost.read_remez(str(GOLDEN / FI))
        fp.approx_md_mother = ost.read_remez(str(GOLDEN / LI))
        fp.approx_li_mother = ost.read_remez(str(GOLDEN / LI))
        self.err_fi = fp.approx_fi_mother.error
        self.err_li = fp.approx_li_mother.error
        set_ptr("fermions_parameters", C.addressof(self.fps))
        ai = ost.AllocSettings.in_dll(lib, "alloc_info")
        ai.NDiffFlavs, ai.NPS_tot, ai.maxApproxOrder, ai.maxNeededShifts = 1, NPS, 19, 19

    def md(self, no_md, t, gauge_scale=2):
        p = ost.MdParam(no_md=no_md, gauge_scale=gauge_scale, t=t)
        mp = ost.MdParam.in_dll(lib, "md_parameters")
        C.memmove(C.addressof(mp), C.addressof(p), C.sizeof(p))
        lib.initialize_md_global_variables(p)

    def update(self, seed, metro=1, acc=0):
        self.conf.put(self.u0)
        lib.ostp_set_rng_seed(seed, 0)
        a = lib.UPDATE_SOLOACC_UNOSTEP_VERSATILE(self.conf.p, 1e-12, 1e-12, 0, acc, metro, 20000)
        ost.check(0, "UPDATE_SOLOACC_UNOSTEP_VERSATILE")
        info = ost.RhmcInfo()
        assert lib.ostp_rhmc_last_info(C.byref(info)) == 0
        return a, info

    def links(self):
        return self.conf.get(self.u0)


def teardown_module(module):
    for name in ("conf_acc_bkp", "aux_conf_acc", "auxbis_conf_acc", "gstout_conf_acc_arr", "momenta", "ipdot_acc",
                 "ferm_phi_acc", "ferm_chi_acc", "ferm_shiftmulti_acc", "k_p_shiftferm", "kloc_r", "kloc_h",
                 "kloc_s", "kloc_p", "fermions_parameters"):
        set_ptr(name, None)
    lib.ostp_set_metropolis_uniform(ost.UNIFORM_FN(), None)


@pytest.fixture(scope="module")
def R():
    return Rhmc()


def test_initial_actions_and_zero_length_trajectory(R):
    R.md(2, 0.0)
    a, info = R.update(99)
    # the reported initial actions, recomputed from the same streams
    du = Dev(R.u0)
    want_g = BETA / 3.0 * (C0_TLSM * lib.calc_plaquette_soloopenacc(du.p, None, None) +
                           C1_TLSM * lib.calc_rettangolo_soloopenacc(du.p, None, None))
    assert abs(info.action_in - want_g) <= 1e-12 * abs(want_g)
    m = Dev(nbytes=8 * 64 * R.o.sizeh)
    lib.ostp_generate_Momenta_gauss(m.p, 99, 0)
    mom = m.get(np.zeros(8 * 8 * R.o.sizeh))
    c = mom.reshape(8, -1)
    sh = R.o.sizeh
    cc = np.ascontiguousarray(c[:, :6 * sh]).view(np.complex128)
    r0, r1 = c[:, 6 * sh:7 * sh], c[:, 7 * sh:]
    want_p = float((r0 ** 2 + r1 ** 2 + r0 * r1).sum() + (np.abs(cc) ** 2).sum())
    assert abs(info.action_mom_in - want_p) <= 1e-12 * want_p
    phi2 = 0.0
    for k in range(NPS):
        v = Dev(nbytes=R.lat.vec3_bytes)
        lib.ostp_generate_vec3_soa_gauss(v.p, 99, 1 + k)
        phi2 += float(np.sum(np.abs(v.get(np.zeros((3, sh), dtype=np.complex128))) ** 2))
    assert abs(info.action_ferm_in - phi2) <= 1e-12 * phi2
    # zero length: only the rational approximations separate the actions
    assert info.action_mom_fin == info.action_mom_in
    assert abs(info.delta_S) <= 5 * phi2 * (2 * R.err_fi + R.err_li), (info.delta_S, phi2)
    assert a == 1 and info.accepted == 1
    assert rel(R.links().view(np.float64), R.u0.view(np.float64)) < 1e-14


def test_delta_s_second_order(R):
    ds = []
    for no_md in (3, 6):
        R.md(no_md, 0.5)
        a, info = R.update(5)
        ds.append(info.delta_S)
        assert info.md_multishift_iterations > 0
    ratio = ds[0] / ds[1]
    assert abs(ratio - 4.0) < 1.0, (ds, ratio)


def test_metropolis_accept_and_reject(R):
    R.md(1, 1.0)

    def run(seed, u, metro=1):
        fn = ost.UNIFORM_FN(lambda ctx: u)
        lib.ostp_set_metropolis_uniform(fn, None)
        try:
            a, info = R.update(seed, metro=metro, acc=3)
        finally:
            lib.ostp_set_metropolis_uniform(ost.UNIFORM_FN(), None)
        return a, info, R.links()

    # a seed whose trajectory raises the action, so that a uniform next to 1 rejects it
    for seed in range(7, 15):
        a, info, links = run(seed, 1.0 - 1e-15)
        if info.delta_S > 0:
            break
    else:
        pytest.fail("no trajectory with dS > 0 among 8 seeds")
    assert info.p1 == pytest.approx(np.exp(-info.delta_S), rel=1e-15) and info.p2 == 1.0 - 1e-15
    assert a == 3 and info.accepted == 0
    assert np.array_equal(links, R.u0)

    a, info, links = run(seed, 0.0)
    assert info.p1 > 0.0, info.delta_S
    assert a == 4 and info.accepted == 1 and info.p2 == 0.0
    a_free, info_free, evolved = run(seed, 0.5, metro=0)
    assert a_free == 4
    assert np.array_equal(links, evolved)
    assert not np.array_equal(evolved, R.u0)
