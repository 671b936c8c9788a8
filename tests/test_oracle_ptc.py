"""Pin the oracle of the gauge-equivariant parallel-transport preconditioner
(oracle/ptc.c): N = w0 + sum_c B_c^dag B_c, B_c = sum_nu w[c][nu] T_nu.

  * closed form: w[0][nu] = +1 (forward), -1 (backward), other channels 0
    gives B_0 = 2 D_oe, hence N = w0 - 4 D_eo D_oe with the (separately
    pinned) oracle Dslash -- fixes every transport's direction, link, phase
    and conjugation;
  * N is Hermitian and <v, N v> >= w0 |v|^2;
  * gauge covariance N(U^g) g r = g N(U) r;
  * the analytic gradient of |N b - x|^2 against central finite differences
    in w0 and in the real and imaginary part of every weight.
"""
import numpy as np
import pytest

from oracle import Oracle
from test_oracle_dslash import random_gauge, random_su3, site_table

C = 12


@pytest.fixture(scope="module")
def sys4():
    rng = np.random.default_rng(3)
    o = Oracle(4, 4, 4, 4)
    u = random_gauge(o, rng)
    ph = o.phases(bf=(0.3, 0.0, -0.7, 1.0, 0.0, 2.0), mu=0.25, charge=1.0)
    w = 0.3 * (rng.standard_normal((C, 8)) + 1j * rng.standard_normal((C, 8)))
    return o, u, ph, rng, w


def test_closed_form_is_the_dslash(sys4):
    o, u, ph, rng, _ = sys4
    w = np.zeros((C, 8), complex)
    w[0, :4], w[0, 4:] = 1.0, -1.0
    r = o.gen_gauss(4)
    z = o.ptc_apply(u, ph, 0.7, w, r)
    ref = 0.7 * r - 4.0 * o.deo(u, o.doe(u, r, ph), ph)
    np.testing.assert_allclose(z, ref, atol=1e-12)
    # and the transports themselves sum to the Dslash
    t = o.ptc_transports(u, ph, r)
    np.testing.assert_allclose(0.5 * (t[:4].sum(0) - t[4:].sum(0)), o.doe(u, r, ph), atol=1e-14)


def test_hermitian_and_positive(sys4):
    o, u, ph, rng, w = sys4
    a, b = o.gen_gauss(5), o.gen_gauss(6)
    w0 = 0.25
    na, nb = o.ptc_apply(u, ph, w0, w, a), o.ptc_apply(u, ph, w0, w, b)
    assert abs(np.vdot(a, nb) - np.conj(np.vdot(b, na))) < 1e-10 * abs(np.vdot(a, nb))
    q = np.vdot(a, na)
    assert abs(q.imag) < 1e-10 * abs(q.real) and q.real >= w0 * np.vdot(a, a).real


def test_gauge_covariance(sys4):
    o, u, ph, rng, w = sys4
    tab = site_table(o)
    g = random_su3(rng, 2 * o.sizeh).reshape(2, o.sizeh, 3, 3)
    L = o.links(u)
    u2 = o.gauge()
    L2 = o.links(u2)
    for d, (i, par) in tab.items():
        for mu in range(4):
            dp = list(d)
            dp[mu] = (dp[mu] + 1) % o.nd[mu]
            j, parp = tab[tuple(dp)]
            L2[2 * mu + par, :, :, i] = g[par, i] @ L[2 * mu + par, :, :, i] @ g[parp, j].conj().T
    r = o.gen_gauss(8)
    gr = np.ascontiguousarray(np.einsum("iab,bi->ai", g[0], r))
    lhs = o.ptc_apply(u2, ph, 0.4, w, gr)
    rhs = np.einsum("iab,bi->ai", g[0], o.ptc_apply(u, ph, 0.4, w, r))
    np.testing.assert_allclose(lhs, rhs, atol=1e-11)


def test_gradient_against_finite_differences(sys4):
    o, u, ph, rng, w = sys4
    b = o.gen_gauss(9)
    x = o.gen_gauss(10)
    w0 = 0.6
    L, g = o.ptc_loss_grad(u, ph, w0, w, b, x)
    h = 1e-6
    checks = [("w0", None)] + [(c, nu) for c, nu in [(0, 0), (3, 5), (11, 7), (6, 2)]]
    for c in checks:
        for part in ((1.0,) if c[0] == "w0" else (1.0, 1j)):
            if c[0] == "w0":
                Lp, _ = o.ptc_loss_grad(u, ph, w0 + h, w, b, x)
                Lm, _ = o.ptc_loss_grad(u, ph, w0 - h, w, b, x)
                ana = g[0]
            else:
                wp, wm = w.copy(), w.copy()
                wp[c] += h * part
                wm[c] -= h * part
                Lp, _ = o.ptc_loss_grad(u, ph, w0, wp, b, x)
                Lm, _ = o.ptc_loss_grad(u, ph, w0, wm, b, x)
                k = 8 * c[0] + c[1]
                ana = g[1 + 2 * k] if part == 1.0 else g[2 + 2 * k]
            fd = (Lp - Lm) / (2 * h)
            assert abs(fd - ana) <= 1e-6 * max(1.0, abs(ana)), (c, part, fd, ana)
