"""Pin the oracle's stout smearing and stout-force chain.

* against the reference's own stouting.c / plaquettes.c / su3_utilities.c,
  compiled from the reference tree into oracle/_ref (to rounding: the oracle
  multiplies 3x3 matrices where the reference unrolls the same products);
* against closed forms: the Cayley-Hamilton exponential equals scipy's expm
  on every branch (small c1 series, near-degenerate theta series, generic),
  smeared links are SU(3), rho = 0 is the identity map.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg

from conftest import ref_lib
from oracle import C_ZERO_TLSM, Oracle, ta_view
from test_oracle_dslash import random_gauge, random_su3

RHO = 0.15


def links(o, u):
    return o.links(u)


def su3_check(o, u):
    L = o.links(u)
    m = np.transpose(L, (0, 3, 1, 2)).reshape(-1, 3, 3).copy()
    # rebuild row 2 as the library does
    m[:, 2, :] = np.conj(np.cross(m[:, 0, :], m[:, 1, :]))
    eye = np.eye(3)
    return np.max(np.abs(m @ np.conj(np.transpose(m, (0, 2, 1))) - eye))


@pytest.fixture(scope="module", params=[(4, 4, 4, 4), (4, 6, 8, 10)])
def field(request):
    rng = np.random.default_rng(11)
    o = Oracle(*request.param)
    u = random_gauge(o, rng)
    return o, u, rng


def test_stout_vs_reference_build(field):
    o, u, _ = field
    ref = ref_lib(o.nd)
    if ref is None:
        pytest.skip("oracle/_ref not built")
    out = o.stout(u, RHO)
    rout = o.gauge()
    ref.ref_stout_isotropic.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]
    ref.ref_stout_isotropic(u.ctypes.data, rout.ctypes.data, RHO)
    A, B = o.links(out)[:, :2], o.links(rout)[:, :2]
    assert np.max(np.abs(A - B)) < 1e-14


def test_staples_and_qa_vs_reference_build(field):
    o, u, _ = field
    ref = ref_lib(o.nd)
    if ref is None:
        pytest.skip("oracle/_ref not built")
    st = o.staples(u)
    q = o.qa(u, st, RHO)
    rst, rq = o.gauge(), o.ta_zero()
    ref.ref_staples_and_qa.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_double]
    ref.ref_staples_and_qa(u.ctypes.data, rst.ctypes.data, rq.ctypes.data, RHO)
    assert np.max(np.abs(o.links(st) - o.links(rst))) < 1e-13
    c, d = ta_view(o, q)
    rc, rd = ta_view(o, rq)
    assert np.max(np.abs(c - rc)) < 1e-14 and np.max(np.abs(d - rd)) < 1e-14


def test_sigma_chain_vs_reference_build(field):
    o, u, rng = field
    ref = ref_lib(o.nd)
    if ref is None:
        pytest.skip("oracle/_ref not built")
    sp = o.gauge()
    S = o.links(sp)
    S[:] = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    s, lam, q = o.sigma_prime_to_sigma(sp, u, RHO)
    rs, rlam, rq = sp.copy(), o.ta_zero(), o.ta_zero()
    ref.ref_sigma_prime_to_sigma.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_double]
    ref.ref_sigma_prime_to_sigma(rs.ctypes.data, rlam.ctypes.data, rq.ctypes.data, u.ctypes.data, RHO)
    c, d = ta_view(o, lam)
    rc, rd = ta_view(o, rlam)
    scale = np.max(np.abs(rc))
    assert np.max(np.abs(c - rc)) < 1e-12 * scale and np.max(np.abs(d - rd)) < 1e-12 * scale
    assert np.max(np.abs(o.links(s) - o.links(rs))) < 1e-12 * np.max(np.abs(o.links(rs)))


def test_unitarize_bitexact_vs_reference_build(field):
    o, u, rng = field
    ref = ref_lib(o.nd)
    if ref is None:
        pytest.skip("oracle/_ref not built")
    v = u.copy()
    L = o.links(v)
    L[:] += 0.05 * (rng.standard_normal(L.shape) + 1j * rng.standard_normal(L.shape))
    a = o.unitarize(v)
    b = v.copy()
    ref.ref_unitarize.argtypes = [ctypes.c_void_p]
    ref.ref_unitarize(b.ctypes.data)
    np.testing.assert_array_equal(o.links(a)[:, :2], o.links(b)[:, :2])
    assert su3_check(o, a) < 1e-14


def test_stout_output_is_su3_and_rho0_is_identity(field):
    o, u, _ = field
    out = o.stout(u, RHO)
    assert su3_check(o, out) < 1e-13
    same = o.stout(u, 0.0)
    assert np.max(np.abs(o.links(same)[:, :2] - o.links(u)[:, :2])) < 1e-15


def tamat_from_matrix(X):
    """QA = traceless anti-Hermitian part of X, as (c01, c02, c12, ic00, ic11)."""
    A = 0.5 * (X - X.conj().T)
    A -= np.trace(A) / 3 * np.eye(3)
    return A, (A[0, 1], A[0, 2], A[1, 2], A[0, 0].imag, A[1, 1].imag)


@pytest.mark.parametrize("scale", [1e-3, 0.02, 0.3, 1.0, 2.5])
def test_ch_exponential_vs_expm(scale):
    """exp(-QA) from the Cayley-Hamilton form (through or_exp_minus_QA_times_conf
    with U = 1) equals scipy.linalg.expm(-QA), on all branches."""
    import oracle as orc

    rng = np.random.default_rng(int(scale * 1000))
    o = Oracle(2, 2, 2, 2)
    n = o.sizeh
    q = o.ta_zero()
    c, d = ta_view(o, q)
    mats = []
    for link in range(8):
        for i in range(n):
            X = scale * (rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3)))
            if (link + i) % 4 == 0:  # near-degenerate spectrum: theta series branch
                X = scale * np.diag([1j, 1j, -2j]) + 1e-5 * X
            A, (c01, c02, c12, i00, i11) = tamat_from_matrix(X)
            c[link, :, i] = (c01, c02, c12)
            d[link, :, i] = (i00, i11)
            mats.append(A)
    u = o.gauge()
    L = o.links(u)
    for k in range(3):
        L[:, k, k, :] = 1.0
    out = o.gauge()
    orc.lib.or_exp_minus_QA_times_conf(o.gp, u.ctypes.data, q.ctypes.data, out.ctypes.data)
    E = o.links(out)
    for idx, A in enumerate(mats):
        link, i = divmod(idx, n)
        ex = scipy.linalg.expm(-A)
        assert np.max(np.abs(E[link, :2, :, i] - ex[:2])) < 1e-14, (scale, link, i)
