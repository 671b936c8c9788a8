"""Pin the oracle's staggered Dslash (restatement of fermion_matrix.c:47-157).

Independent checks that a wrong index, sign, conjugation, phase or factor
would fail:
  * anti-Hermiticity  D_oe = -D_eo^dagger on random SU(3) links and phases;
  * gauge covariance  D(U^g) g psi = g D(U) psi;
  * free field: the spectrum of m^2 - D_eo D_oe with U = 1 and antiperiodic
    time is m^2 + sum_mu sin^2 p_mu, p_t = (2n+1) pi / L_t, with the staggered
    multiplicities;
  * the FP32 twin agrees with FP64 to single precision.
"""
import itertools
import math

import numpy as np
import pytest

import oracle as orc
from oracle import Oracle


def random_su3(rng, n):
    """n Haar-ish SU(3) matrices (QR of complex Gaussians, det fixed to 1)."""
    z = (rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    q = q / (det ** (1.0 / 3.0))[:, None, None]
    return q


def random_gauge(o, rng):
    u = o.gauge()
    L = o.links(u)
    m = random_su3(rng, 8 * o.sizeh).reshape(8, o.sizeh, 3, 3)
    L[:] = np.transpose(m, (0, 2, 3, 1))
    return u


def dense(o, apply, u, ph):
    n = 3 * o.sizeh
    M = np.zeros((n, n), dtype=np.complex128)
    for j in range(n):
        e = o.vec3()
        e.reshape(-1)[j] = 1.0
        M[:, j] = apply(u, e, ph).reshape(-1)
    return M


@pytest.fixture(scope="module")
def small():
    rng = np.random.default_rng(7)
    o = Oracle(4, 4, 4, 4)
    u = random_gauge(o, rng)
    ph = o.phases(bf=(0.3, 0.0, -0.7, 1.0, 0.0, 2.0), mu=0.25, charge=1.0)
    return o, u, ph, rng


def test_anti_hermitian(small):
    o, u, ph, _ = small
    Deo = dense(o, o.deo, u, ph)
    Doe = dense(o, o.doe, u, ph)
    np.testing.assert_allclose(Doe, -Deo.conj().T, atol=1e-14)
    # hence m^2 - Deo Doe is Hermitian positive definite
    M = 0.01 * np.eye(Deo.shape[0]) - Deo @ Doe
    np.testing.assert_allclose(M, M.conj().T, atol=1e-13)
    assert np.linalg.eigvalsh(M).min() >= 0.01 - 1e-12


def site_table(o):
    tab = {}
    for d3, d2, d1, d0 in itertools.product(*(range(n) for n in o.nd[::-1])):
        d = (d0, d1, d2, d3)
        tab[d] = (orc.lib.or_snum(o.gp, *d), sum(d) % 2)
    return tab


def test_gauge_covariance(small):
    o, u, ph, rng = small
    tab = site_table(o)
    g = random_su3(rng, 2 * o.sizeh).reshape(2, o.sizeh, 3, 3)  # g[parity, idxh]
    L = o.links(u)
    u2 = o.gauge()
    L2 = o.links(u2)
    for d, (i, par) in tab.items():
        for mu in range(4):
            dp = list(d)
            dp[mu] = (dp[mu] + 1) % o.nd[mu]
            j, parp = tab[tuple(dp)]
            U = L[2 * mu + par, :, :, i]
            L2[2 * mu + par, :, :, i] = g[par, i] @ U @ g[parp, j].conj().T
    psi_o = o.gen_gauss(11)
    gpsi = np.einsum("iab,bi->ai", g[1], psi_o)
    lhs = o.deo(u2, np.ascontiguousarray(gpsi), ph)
    rhs = np.einsum("iab,bi->ai", g[0], o.deo(u, psi_o, ph))
    np.testing.assert_allclose(lhs, rhs, atol=1e-13)
    psi_e = o.gen_gauss(12)
    gpsi = np.einsum("iab,bi->ai", g[0], psi_e)
    lhs = o.doe(u2, np.ascontiguousarray(gpsi), ph)
    rhs = np.einsum("iab,bi->ai", g[1], o.doe(u, psi_e, ph))
    np.testing.assert_allclose(lhs, rhs, atol=1e-13)


@pytest.mark.parametrize("shape", [(4, 4, 4, 4), (4, 2, 2, 6)])
def test_free_field_spectrum(shape):
    o = Oracle(*shape)
    u = o.gauge()
    L = o.links(u)
    for c in range(3):
        L[:, c, c, :] = 1.0
    ph = o.phases()
    m = 0.05
    Deo = dense(o, o.deo, u, ph)
    Doe = dense(o, o.doe, u, ph)
    ev = np.sort(np.linalg.eigvalsh(m * m * np.eye(Deo.shape[0]) - Deo @ Doe))
    # momenta of the full lattice; the even-site operator carries each
    # sum_mu sin^2 p_mu value with multiplicity 3 colours x 1/2 (even sites)
    nx, ny, nz, nt = shape
    vals = []
    for kx, ky, kz, kt in itertools.product(range(nx), range(ny), range(nz), range(nt)):
        p = [2 * math.pi * kx / nx, 2 * math.pi * ky / ny, 2 * math.pi * kz / nz, (2 * kt + 1) * math.pi / nt]
        vals.append(m * m + sum(math.sin(q) ** 2 for q in p))
    vals = np.sort(np.repeat(np.array(vals), 3))  # colours
    # staggered: each eigenvalue appears on both parities; keep half
    np.testing.assert_allclose(ev, vals[::2], atol=1e-12)


def test_fp32_twin_matches_fp64(small):
    o, u, ph, _ = small
    uf = o.to_float_gauge(u)
    phf = ph.astype(np.float32)
    psi = o.gen_gauss(5)
    ref = o.deo(u, psi, ph)
    got = o.deo_f(uf, psi.astype(np.complex64), phf)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 5e-7
    ref = o.doe(u, psi, ph)
    got = o.doe_f(uf, psi.astype(np.complex64), phf)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 5e-7


def test_mdagm_composition(small):
    o, u, ph, _ = small
    psi = o.gen_gauss(3)
    m, sh = 0.1, 0.37
    out = o.mdagm(u, psi, ph, m, sh)
    expect = (m * m + sh) * psi - o.deo(u, o.doe(u, psi, ph), ph)
    np.testing.assert_allclose(out, expect, atol=1e-14)
