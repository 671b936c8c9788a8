"""The fermion force is the derivative of the fermion action.

fermion_force_soloopenacc (fermion_force.c:166-358) returns, per link, the
traceless anti-Hermitian ipdot_f whose momentum step is P += delta ipdot_f
with the same delta sign convention as the gauge force (md_parameters.c:30-41,
su3_utilities.c:241-270): for the smeared staggered action

    S_f(U) = sum_flavours sum_pseudofermions < phi, r(M^+M[V(U)]) phi >,
    V = stout^2(U) (rho = 0.15),  r = a0 + sum_k a_k / (x + b_k),

and a left variation U_l(x) -> exp(i eps T_a) U_l(x) along an su(3) generator,

    dS_f/d eps = - Re tr(i T_a ipdot_f(l, x)).

S_f is evaluated here without any solver: M^+M is built as a dense matrix on a
4^4 lattice from the oracle's Dslash (4^4: 384 x 384) and r is applied through
its eigen-decomposition, so the check is independent of the CG-M, the outer
products, the phase multiplication and the stout force chain it pins.  Cases:
no background field (every phase +-1), and two flavours with electric and
magnetic fields and an imaginary chemical potential (complex phases, different
charges), where a conjugation or sign slip in the phase multiply or in the
outer products would show.

The CPU test checks the oracle composition; the GPU twin checks the library's
fermion_force_soloopenacc against the same derivatives and against the oracle
composition (1e-7, the solver tolerance).
"""
import ctypes as C
import pathlib

import numpy as np
import pytest

import openstaple_amd as ost
from oracle import Oracle, ta_view

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
RHO = 0.15
STEPS = 2
SHAPE = (4, 4, 4, 4)

# (mass, background field (ex, ey, ez, bx, by, bz), imaginary mu, charge)
CASES = {
    "no_field": [(0.3, (0, 0, 0, 0, 0, 0), 0.0, 0.0)],
    "em_field_imag_mu": [(0.3, (0.7, -1.3, 0.4, 1.0, 3.0, -2.0), 0.25, 2.0 / 3.0),
                         (0.45, (0.7, -1.3, 0.4, 1.0, 3.0, -2.0), 0.25, -1.0 / 3.0)],
}


def approx_md9():
    ra = ost.read_remez(str(GOLDEN / "approx_-1_over_4_order_9_mloglm_6.4.REMEZ"))
    lmax = 32.0
    eps = lmax ** (ra.exponent_num / ra.exponent_den)
    ra.RA_a0 *= eps
    for k in range(ra.approx_order):
        ra.RA_a[k] *= lmax * eps
        ra.RA_b[k] *= lmax
    return ra


def gell_mann():
    l = np.zeros((8, 3, 3), dtype=np.complex128)
    l[0][0, 1] = l[0][1, 0] = 1
    l[1][0, 1], l[1][1, 0] = -1j, 1j
    l[2][0, 0], l[2][1, 1] = 1, -1
    l[3][0, 2] = l[3][2, 0] = 1
    l[4][0, 2], l[4][2, 0] = -1j, 1j
    l[5][1, 2] = l[5][2, 1] = 1
    l[6][1, 2], l[6][2, 1] = -1j, 1j
    l[7] = np.diag([1, 1, -2]) / np.sqrt(3)
    return l / 2


def exp_i(h):
    """exp(i h) of a Hermitian 3x3 matrix."""
    w, v = np.linalg.eigh(h)
    return (v * np.exp(1j * w)) @ v.conj().T


class Setup:
    def __init__(self, case):
        self.o = o = Oracle(*SHAPE)
        self.u = o.gen_cold(0.5, 11)
        self.ra = approx_md9()
        self.flavours = []
        for f, (mass, bf, mu, charge) in enumerate(CASES[case]):
            ph = o.phases(bf=bf, mu=mu, charge=charge)
            phi = o.gen_gauss(300 + f)
            self.flavours.append((mass, ph, phi))
        if case != "no_field":  # the phases really are complex
            assert np.mean(np.abs(np.sin(self.flavours[0][1]))) > 0.3

    def smear(self, u):
        levels = [u]
        for _ in range(STEPS):
            levels.append(self.o.stout(levels[-1], RHO))
        return levels

    def dense_mdagm(self, v, ph, mass):
        o = self.o
        n = 3 * o.sizeh
        A = np.zeros((n, n), dtype=np.complex128)
        e = o.vec3()
        for j in range(n):
            e[:] = 0
            e.reshape(-1)[j] = 1.0
            col = mass * mass * e - o.deo(v, o.doe(v, e, ph), ph)
            A[:, j] = col.reshape(-1)
        return A

    def action(self, u):
        v = self.smear(u)[-1]
        ra = self.ra
        tot = 0.0
        for mass, ph, phi in self.flavours:
            A = self.dense_mdagm(v, ph, mass)
            assert np.abs(A - A.conj().T).max() < 1e-12 * np.abs(A).max()
            w, Q = np.linalg.eigh(0.5 * (A + A.conj().T))
            assert w.min() > 0 and w.max() < 32.0
            r = ra.RA_a0 + sum(ra.RA_a[k] / (w + ra.RA_b[k]) for k in range(ra.approx_order))
            c = Q.conj().T @ phi.reshape(-1)
            tot += float(np.sum(r * np.abs(c) ** 2))
        return tot

    def oracle_force(self):
        """The oracle composed as fermion_force.c:166-358 composes the reference."""
        o, ra = self.o, self.ra
        levels = self.smear(self.u)
        v = levels[-1]
        shifts = [ra.RA_b[k] for k in range(ra.approx_order)]
        gl3 = o.gauge()
        for mass, ph, phi in self.flavours:
            taux = o.gauge()
            st, cg, xs = o.multishift(v, ph, mass, shifts, np.ascontiguousarray(phi), 1e-12)
            assert st == 1
            for k in range(ra.approx_order):
                xk = np.ascontiguousarray(xs[k])
                o.direct_product(xk, o.doe(v, xk, ph), taux, ra.RA_a[k])
            o.backfield_times_force(ph, taux, gl3)
        for level in range(STEPS, 0, -1):
            gl3, _, _ = o.sigma_prime_to_sigma(gl3, levels[level - 1], RHO)
        ip = o.ta_zero()
        o.conf_times_force_ta(self.u, gl3, ip)
        return ip

    def check_derivatives(self, ip, npoints=6, seed=0):
        """-Re tr(i T_a ipdot) against central differences of S_f."""
        o = self.o
        c, d = ta_view(o, ip)
        T = gell_mann()
        rng = np.random.default_rng(seed)
        links = o.links(self.u)
        eps = 1e-4
        worst = 0.0
        sizes = []
        for _ in range(npoints):
            l, i, a = int(rng.integers(8)), int(rng.integers(o.sizeh)), int(rng.integers(8))
            A = np.array([[1j * d[l, 0, i], c[l, 0, i], c[l, 1, i]],
                          [-np.conj(c[l, 0, i]), 1j * d[l, 1, i], c[l, 2, i]],
                          [-np.conj(c[l, 1, i]), -np.conj(c[l, 2, i]), -1j * (d[l, 0, i] + d[l, 1, i])]])
            analytic = -np.real(np.trace(1j * T[a] @ A))
            U = links[l, :, :, i].copy()
            got = []
            for s in (1, -1):
                u2 = self.u.copy()
                o.links(u2)[l, :, :, i] = exp_i(s * eps * T[a]) @ U
                got.append(self.action(u2))
            fd = (got[0] - got[1]) / (2 * eps)
            sizes.append(abs(fd))
            err = abs(fd - analytic) / max(abs(fd), 1e-3)
            worst = max(worst, err)
            assert abs(fd - analytic) <= 1e-6 * max(1.0, abs(fd)), (l, i, a, fd, analytic)
        assert max(sizes) > 1e-2  # the derivatives probed are not trivially zero
        return worst


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_force_is_action_derivative(case):
    S = Setup(case)
    S.check_derivatives(S.oracle_force())


def test_oracle_force_sign_slip_is_caught():
    """The check fails for a force whose phase multiply is conjugated (the
    slip it is there to catch): the same derivative test on the oracle with
    conj(e^{i theta}) in the phase step."""
    S = Setup("em_field_imag_mu")
    o, ra = S.o, S.ra
    levels = S.smear(S.u)
    v = levels[-1]
    shifts = [ra.RA_b[k] for k in range(ra.approx_order)]
    gl3 = o.gauge()
    for mass, ph, phi in S.flavours:
        taux = o.gauge()
        _, _, xs = o.multishift(v, ph, mass, shifts, np.ascontiguousarray(phi), 1e-12)
        for k in range(ra.approx_order):
            xk = np.ascontiguousarray(xs[k])
            o.direct_product(xk, o.doe(v, xk, ph), taux, ra.RA_a[k])
        o.backfield_times_force(np.ascontiguousarray(-ph), taux, gl3)  # the slip: e^{-i theta}
    for level in range(STEPS, 0, -1):
        gl3, _, _ = o.sigma_prime_to_sigma(gl3, levels[level - 1], RHO)
    ip = o.ta_zero()
    o.conf_times_force_ta(S.u, gl3, ip)
    with pytest.raises(AssertionError):
        S.check_derivatives(ip, npoints=4)


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_fermion_force_is_action_derivative(case):
    from openstaple_amd import lib
    from gpu_util import Dev, init_lattice

    S = Setup(case)
    o, ra = S.o, S.ra
    init_lattice(SHAPE)
    lib.ostp_set_stout(RHO, STEPS, 0.0, 0, 0, 1)
    lib.ostp_set_inverter_tricks(0, 0, 0.1, 10000)
    lat = ost.Lattice(*SHAPE)
    order = ra.approx_order
    nf = len(S.flavours)
    du = Dev(S.u)
    dphs = [Dev(ph) for _, ph, _ in S.flavours]
    dphi = Dev(np.stack([phi for _, _, phi in S.flavours]))
    darr = Dev(nbytes=STEPS * S.u.nbytes)
    dgl3, dtaux = Dev(nbytes=S.u.nbytes), Dev(nbytes=S.u.nbytes)
    ipref = S.oracle_force()
    dip = Dev(nbytes=ipref.nbytes)
    dxs, dsf = Dev(nbytes=order * lat.vec3_bytes), Dev(nbytes=order * lat.vec3_bytes)
    d64 = [Dev(nbytes=lat.vec3_bytes) for _ in range(4)]
    fps = (ost.FermParam * nf)()
    for f, (mass, _, _) in enumerate(S.flavours):
        fps[f].ferm_mass = mass
        fps[f].phases = dphs[f].p
        fps[f].number_of_ps = 1
        fps[f].index_of_the_first_ps = f
        fps[f].approx_md = ra
    ipk = ost.InverterPackage()
    ipk.u = du.p
    ipk.ferm_shift_temp, ipk.nshifts = dsf.p, order
    ipk.loc_r, ipk.loc_h, ipk.loc_s, ipk.loc_p = (d.p for d in d64)
    lib.fermion_force_soloopenacc(du.p, darr.p, dgl3.p, dip.p, fps, nf, dphi.p, 1e-12, dtaux.p, dxs.p, ipk, 10000)
    ost.check(0, "fermion_force_soloopenacc")
    ip = dip.get(ipref)
    c, d = ta_view(o, ip)
    rc, rd = ta_view(o, ipref)
    scale = max(np.abs(rc).max(), np.abs(rd).max())
    assert np.abs(c - rc).max() <= 1e-7 * scale and np.abs(d - rd).max() <= 1e-7 * scale
    S.check_derivatives(ip, seed=1)
