"""Pin the oracle's solvers (ker_invert_openacc, inverter_mixed_precision,
multishift_invert, recombine_shifted_vec3_to_vec3) against dense linear
algebra on a 4^4 lattice, and the rational approximations against the error
their .REMEZ headers state."""
import pathlib

import numpy as np
import pytest

from oracle import Oracle, read_remez
from test_oracle_dslash import dense, random_gauge

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def sys44():
    rng = np.random.default_rng(3)
    o = Oracle(4, 4, 4, 4)
    u = random_gauge(o, rng)
    ph = o.phases()
    Deo = dense(o, o.deo, u, ph)
    Doe = dense(o, o.doe, u, ph)
    K = -Deo @ Doe
    b = o.gen_gauss(21)
    return o, u, ph, K, b


def test_cg_vs_dense(sys44):
    o, u, ph, K, b = sys44
    m = 0.2
    st, cg, x = o.cg(u, ph, m, b, 1e-10)
    assert st == 1 and 0 < cg < 384
    xd = np.linalg.solve(m * m * np.eye(K.shape[0]) + K, b.reshape(-1)).reshape(3, -1)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    r = b - o.mdagm(u, x, ph, m)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10


def test_cg_restarts_and_shift(sys44):
    o, u, ph, K, b = sys44
    m, sh = 0.1, 0.05
    st, cg, x = o.cg(u, ph, m, b, 1e-9, shift=sh, restart=7)
    assert st == 1
    xd = np.linalg.solve((m * m + sh) * np.eye(K.shape[0]) + K, b.reshape(-1)).reshape(3, -1)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-7


def test_cg_max_cg_failure(sys44):
    o, u, ph, K, b = sys44
    st, cg, x = o.cg(u, ph, 0.01, b, 1e-12, max_cg=5, restart=5)
    assert st == 0 and cg == 5


def test_cg_warm_start_converges_immediately(sys44):
    o, u, ph, K, b = sys44
    m = 0.2
    _, _, x = o.cg(u, ph, m, b, 1e-12)
    st, cg, x2 = o.cg(u, ph, m, b, 1e-9, x0=x)
    assert st == 1 and cg == 1


def test_mixed_precision(sys44):
    o, u, ph, K, b = sys44
    m = 0.1
    uf = o.to_float_gauge(u)
    phf = ph.astype(np.float32)
    st, cg, x = o.mixed(u, uf, ph, phf, m, b, 1e-10)
    assert st == 1
    r = b - o.mdagm(u, x, ph, m)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10


def test_multishift_vs_dense(sys44):
    o, u, ph, K, b = sys44
    m = 0.1
    shifts = [0.0, 1e-3, 0.05, 0.7, 3.0]
    st, cg, xs = o.multishift(u, ph, m, shifts, b, 1e-10)
    assert st == 1
    n = K.shape[0]
    for k, s in enumerate(shifts):
        xd = np.linalg.solve((m * m + s) * np.eye(n) + K, b.reshape(-1)).reshape(3, -1)
        rel = np.linalg.norm(xs[k] - xd) / np.linalg.norm(xd)
        assert rel < 1e-8, (k, rel)


def test_recombine_is_rational_function_of_the_operator(sys44):
    o, u, ph, K, b = sys44
    m = 0.1
    ra = read_remez(GOLDEN / "approx_-1_over_4_order_19_mloglm_6.4.REMEZ")
    order = ra.approx_order
    shifts = [ra.RA_b[k] for k in range(order)]
    a = [ra.RA_a[k] for k in range(order)]
    # scale into the operator's spectral range like rescale_rational_approximation
    A = m * m * np.eye(K.shape[0]) + K
    ev, V = np.linalg.eigh(A)
    st, cg, xs = o.multishift(u, ph, m, shifts, b, 1e-11)
    chi = o.recombine(xs, b, ra.RA_a0, a)
    rvals = ra.RA_a0 + sum(a[k] / (ev + shifts[k]) for k in range(order))
    expect = (V @ (rvals * (V.conj().T @ b.reshape(-1)))).reshape(3, -1)
    assert np.linalg.norm(chi - expect) / np.linalg.norm(expect) < 1e-9


def _remez_kind(path):
    txt = path.read_text()
    if "Order:" not in txt:
        return "headerless"
    err_line = [l for l in txt.splitlines() if l.startswith("Error:")][0]
    return "full" if "e" in err_line.split(":")[1] else "legacy"


REMEZ = sorted(GOLDEN.glob("*.REMEZ"))


@pytest.mark.parametrize("path", [p for p in REMEZ if _remez_kind(p) == "full"], ids=lambda p: p.name)
def test_remez_files_meet_their_stated_error(path):
    """max |r(x) / x^(num/den) - 1| on [lambda_min, lambda_max] is the error the
    file header states (Remez equioscillation: the maximum is attained)."""
    ra = read_remez(path)
    power = ra.exponent_num / ra.exponent_den
    x = np.geomspace(ra.lambda_min, ra.lambda_max, 20001)
    r = ra.RA_a0 + sum(ra.RA_a[k] / (x + ra.RA_b[k]) for k in range(ra.approx_order))
    err = np.max(np.abs(r / x ** power - 1.0))
    assert err <= ra.error * 1.05, (err, ra.error)
    assert err >= ra.error * 0.5


@pytest.mark.parametrize("path", [p for p in REMEZ if _remez_kind(p) == "legacy"], ids=lambda p: p.name)
def test_remez_legacy_files_parse(path):
    """Older files print coefficients with %f: they parse and approximate to ~1e-5."""
    ra = read_remez(path)
    power = ra.exponent_num / ra.exponent_den
    x = np.geomspace(ra.lambda_min, ra.lambda_max, 2001)
    r = ra.RA_a0 + sum(ra.RA_a[k] / (x + ra.RA_b[k]) for k in range(ra.approx_order))
    assert np.max(np.abs(r / x ** power - 1.0)) < 1e-4


@pytest.mark.parametrize("path", [p for p in REMEZ if _remez_kind(p) == "headerless"], ids=lambda p: p.name)
def test_remez_headerless_files_are_rejected(path):
    """rationalapprox_read_custom_nomefile's CHECKREAD rejects files without the
    Order/Lambda header (rationalapprox.c:9-13,93-99)."""
    with pytest.raises(ValueError):
        read_remez(path)


@pytest.fixture(scope="module")
def dense44():
    """M = m^2 - Deo Doe on a 4^4 expH field as a dense 384 x 384 matrix."""
    o = Oracle(4, 4, 4, 4)
    u = o.gen_expH(0.3, 1)
    ph = o.phases()
    mass = 0.05
    n = 3 * o.sizeh
    M = np.zeros((n, n), dtype=np.complex128)
    for j in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[j] = 1.0
        M[:, j] = o.mdagm(u, e.reshape(3, -1), ph, mass).ravel()
    w, V = np.linalg.eigh((M + M.conj().T) / 2)
    return o, u, ph, mass, M, w, V


def test_power_method_max_eigenvalue(dense44):
    """ker_find_max_eigenvalue_openacc restated: from a Gaussian start the
    estimate approaches lambda_max from below (||M r|| <= lambda_max for unit r)
    and stops within the 1e-5 change criterion; from the exact top eigenvector
    it returns lambda_max after two iterations (the second sees no change)."""
    o, u, ph, mass, M, w, V = dense44
    assert np.max(np.abs(M - M.conj().T)) < 1e-14
    b = o.gen_gauss(3)
    v, it = o.find_max_eigenvalue(u, ph, mass, b)
    assert w[-1] * (1 - 1e-3) <= v <= w[-1] * (1 + 1e-12) and it > 10
    top = np.ascontiguousarray(V[:, -1].reshape(3, -1))
    v2, it2 = o.find_max_eigenvalue(u, ph, mass, top)
    assert abs(v2 - w[-1]) < 1e-12 * w[-1] and it2 == 2
    # the mass term is in the operator: the same with m = 0.3 shifts lambda_max by 0.3^2 - 0.05^2
    v3, _ = o.find_max_eigenvalue(u, ph, 0.3, top)
    assert abs(v3 - (w[-1] + 0.09 - 0.0025)) < 1e-12 * w[-1]


def test_power_method_min_formula(dense44):
    """ker_find_min_eigenvalue_openacc restated: its operator is
    (m^2 + (max - m^2) - m^2) - Deo Doe = M + (max - 2 m^2), whose top eigenvalue
    is lambda_max + max - 2 m^2, so it returns 2 m^2 - lambda_max (the caller,
    find_min_max_eigenvalue_soloopenacc, uses m^2 instead)."""
    o, u, ph, mass, M, w, V = dense44
    top = np.ascontiguousarray(V[:, -1].reshape(3, -1))
    vmax = w[-1]
    vmin, it = o.find_min_eigenvalue(u, ph, mass, top, vmax)
    assert abs(vmin - (2 * mass * mass - w[-1])) < 1e-11 * w[-1] and it == 2


def test_eo_inversion_condensate_vs_dense_inverse():
    """The chiral-condensate estimator eta^dag M^-1 eta (ferm_meas.c:282-307)
    through the oracle's eo_inversion equals the one from the dense inverse of
    the full staggered matrix M = m + D (D = Deo on odd -> even, Doe on even
    -> odd) at 4^4; averaged over Z2 noise it estimates tr M^-1."""
    o = Oracle(4, 4, 4, 4)
    u = o.gen_expH(0.3, 1)
    ph = o.phases()
    mass = 0.1
    n = 3 * o.sizeh
    D = np.zeros((2 * n, 2 * n), dtype=np.complex128)
    for j in range(n):
        e = np.zeros((3, o.sizeh), dtype=np.complex128)
        e.reshape(-1)[j] = 1.0
        D[:n, n + j] = o.deo(u, e, ph).ravel()   # odd input j -> even output
        D[n:, j] = o.doe(u, e, ph).ravel()       # even input j -> odd output
    M = mass * np.eye(2 * n) + D
    Minv = np.linalg.inv(M)
    ests = []
    for k in range(64):
        be, bo = o.gen_z2(5, 2 * k), o.gen_z2(5, 2 * k + 1)
        st, _, xe, xo = o.eo_inversion(u, ph, mass, be, bo, 1e-13)
        eta = np.concatenate([be.ravel(), bo.ravel()])
        est = np.vdot(be, xe) + np.vdot(bo, xo)
        exact = np.vdot(eta, Minv @ eta)
        assert st == 1 and abs(est - exact) <= 1e-9 * abs(exact)
        ests.append(est)
    ests = np.array(ests)
    tr = np.trace(Minv)
    # Z2 noise: unbiased estimator of the trace; 64 samples, 5 sigma
    assert abs(ests.mean() - tr) <= 5 * ests.std() / np.sqrt(len(ests)), (ests.mean(), tr)
    # D is anti-Hermitian, so Re tr M^-1 = sum m / (m^2 + |lambda|^2) > 0 and Im tr M^-1 = 0
    assert tr.real > 0 and abs(tr.imag) < 1e-10 * abs(tr)
