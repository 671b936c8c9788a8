"""Stopping rules of the CG and CG-M inverters against textbook residual
histories on a dense 4^4 system (384 x 384, built from the oracle's Dslash).

The rules (inverter_full.c:17,73-105; inverter_multishift_full.c:18,161-176):
* CG iterates until sqrt(|r|^2 / |b|^2) <= 0.95 res (the inner test, on the
  recursive residual), restarting from the true residual every
  `restartingEvery` iterations; the outer loop (checked only at a restart or
  an inner exit) continues while the recursive residual is above res and
  fewer than max_cg iterations were done -- so max_cg is not checked inside
  an inner cycle;
* CG-M turns shift k off at iteration j + 1 when its shifted residual at the
  start of that iteration, |zeta_j^k| |r_j| / |b|, is below 0.95 res, and
  stops when every shift is off.

The residual histories here come from numpy, not from the solvers under test:
CG as in Saad, Iterative Methods for Sparse Linear Systems, alg. 6.18, on the
dense matrix; the shifted residual norms of CG-M from the Lanczos relation
(Saad eq. 6.103): the CG residual polynomial pi_j has the Ritz values of the
tridiagonal T_j as roots, and the shifted residual is r_j / pi_j(-sigma).
`res` is chosen from the history so that a residual falls between 0.95 res
and res: with a safety factor of 1.0 instead of 0.95 the solvers would stop
one or more iterations earlier, and the tests would fail.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import Oracle
from test_oracle_dslash import dense, random_gauge

SAFETY = 0.95
ITERS = 400  # iterations of the histories (CG at m = 0.1 reaches 1e-8 in 232)
MARGIN = 1e-5  # relative distance a residual keeps from a threshold it is compared with


@pytest.fixture(scope="module")
def system():
    rng = np.random.default_rng(5)
    o = Oracle(4, 4, 4, 4)
    u = random_gauge(o, rng)
    ph = o.phases()
    K = -dense(o, o.deo, u, ph) @ dense(o, o.doe, u, ph)
    b = o.gen_gauss(44)
    return o, u, ph, K, b


def cg_history(A, b, x0, iters, restart_every=10 ** 9):
    """Recursive residual norms sqrt(|r_k|^2/|b|^2), k = 1..iters, of CG with
    restarts from the true residual every restart_every iterations."""
    x = x0.copy()
    bn = np.vdot(b, b).real
    hist = []
    while len(hist) < iters:
        r = b - A @ x
        p = r.copy()
        rr = np.vdot(r, r).real
        for _ in range(min(restart_every, iters - len(hist))):
            s = A @ p
            a = rr / np.vdot(p, s).real
            x = x + a * p
            r = r - a * s
            rr_new = np.vdot(r, r).real
            p = r + (rr_new / rr) * p
            rr = rr_new
            hist.append(np.sqrt(rr / bn))
    return np.array(hist)


def cg_count(hist, res, restart_every, max_cg):
    """Iterations the reference's rule takes on a history."""
    k = 0
    while True:
        cr = 0
        while True:
            k += 1
            cr += 1
            if not (hist[k - 1] > res * SAFETY and cr < restart_every):
                break
        if not (hist[k - 1] > res and k < max_cg):
            return k


def separated(vals, thr):
    return np.all(np.abs(np.asarray(vals) / thr - 1.0) > MARGIN)


def pick_res(hist, below=1e-8):
    """res near `below` with some hist[k] in (0.95 res, res] and every value
    near both thresholds kept MARGIN away; returns (res, the k with res*0.95
    first reached, the k with res first reached)."""
    assert hist.min() < below
    for k in range(int(np.argmax(hist < below)), len(hist) - 5):
        res = hist[k] / 0.975
        if not (separated(hist, res) and separated(hist, SAFETY * res)):
            continue
        first95 = int(np.argmax(hist <= SAFETY * res)) + 1
        first100 = int(np.argmax(hist <= res)) + 1
        if hist.min() <= SAFETY * res and first100 < first95:
            return res, first95, first100
    raise AssertionError("no usable threshold")


def test_cg_stops_at_the_first_residual_below_the_safety_margin(system):
    o, u, ph, K, b = system
    m = 0.1
    A = m * m * np.eye(K.shape[0]) + K
    bv = b.reshape(-1)
    hist = cg_history(A, bv, np.zeros_like(bv), ITERS)
    res, first95, first100 = pick_res(hist)
    assert cg_count(hist, res, 10 ** 9, 10000) == first95
    st, cg, x = o.cg(u, ph, m, b, res)
    assert cg == first95 and st == 1
    assert cg > first100  # a safety factor of 1.0 would have stopped earlier


def test_cg_restart_and_max_cg_rules(system):
    o, u, ph, K, b = system
    m = 0.1
    A = m * m * np.eye(K.shape[0]) + K
    bv = b.reshape(-1)
    R = 40
    hist = cg_history(A, bv, np.zeros_like(bv), ITERS, restart_every=R)
    res, first95, _ = pick_res(hist)
    n = cg_count(hist, res, R, 10000)
    st, cg, _ = o.cg(u, ph, m, b, res, restart=R)
    assert cg == n and st == 1
    # with restarts the count differs from the unrestarted one (the rule is exercised)
    assert n > cg_count(cg_history(A, bv, np.zeros_like(bv), ITERS), res, 10 ** 9, 10000)
    # max_cg is tested only when an inner cycle ends: without restarts CG runs to convergence
    st, cg, _ = o.cg(u, ph, m, b, res, max_cg=7)
    assert cg == cg_count(cg_history(A, bv, np.zeros_like(bv), ITERS), res, 10 ** 9, 7) > 7 and st == 1
    # with restarts every R it stops at the first restart boundary at or beyond max_cg
    st, cg, _ = o.cg(u, ph, m, b, 1e-14, max_cg=60, restart=R)
    assert cg == 80 and st == 0


def shifted_residual_history(A, b, shifts, iters):
    """|r_j^sigma| / |b| for j = 0..iters of the CG-M iterates, from the
    Ritz values of the CG-Lanczos tridiagonal matrix."""
    bv = b
    bn = np.sqrt(np.vdot(bv, bv).real)
    x = np.zeros_like(bv)
    r = bv.copy()
    p = r.copy()
    rr = np.vdot(r, r).real
    alphas, betas, rnorm = [], [], [1.0]
    for _ in range(iters):
        s = A @ p
        a = rr / np.vdot(p, s).real
        r = r - a * s
        rr_new = np.vdot(r, r).real
        beta = rr_new / rr
        p = r + beta * p
        rr = rr_new
        alphas.append(a)
        betas.append(beta)
        rnorm.append(np.sqrt(rr) / bn)
    out = np.zeros((len(shifts), iters + 1))
    out[:, 0] = 1.0
    for j in range(1, iters + 1):
        T = np.zeros((j, j))
        for i in range(j):
            T[i, i] = 1.0 / alphas[i] + (betas[i - 1] / alphas[i - 1] if i > 0 else 0.0)
            if i + 1 < j:
                T[i, i + 1] = T[i + 1, i] = np.sqrt(betas[i]) / alphas[i]
        theta = np.linalg.eigvalsh(T)
        for k, sg in enumerate(shifts):
            out[k, j] = rnorm[j] / abs(np.prod(1.0 + sg / theta))
    return out


def cgm_count(shist, res):
    """Iterations of the per-shift flag rule: shift k is off after iteration
    j + 1 for the first j with its shifted residual below 0.95 res."""
    ends = []
    for h in shist:
        j = int(np.argmax(h < SAFETY * res))
        assert h[j] < SAFETY * res
        ends.append(j + 1)
    return max(ends), ends


def test_cgm_per_shift_flags(system):
    o, u, ph, K, b = system
    m = 0.1
    A = m * m * np.eye(K.shape[0]) + K
    shifts = [0.0, 0.004, 0.05, 0.6, 4.0]
    sh = shifted_residual_history(A, b.reshape(-1), shifts, ITERS)
    slow = sh[0]
    # res near 1e-8 with the slowest shift's history crossing 0.95 res one or more iterations after res
    for k in range(int(np.argmax(slow < 1e-8)), ITERS - 5):
        res = slow[k] / 0.975
        if separated(sh.ravel(), res * SAFETY) and np.argmax(slow < SAFETY * res) > np.argmax(slow < res):
            break
    else:
        raise AssertionError("no usable threshold")
    n, ends = cgm_count(sh, res)
    assert len(set(ends)) > 2  # shifts stop at different iterations
    st, cg, xs = o.multishift(u, ph, m, shifts, b, res)
    assert cg == n and st == 1
    # every shift's solution reaches the tolerance
    for k, sg in enumerate(shifts):
        r = b.reshape(-1) - (A + sg * np.eye(A.shape[0])) @ xs[k].reshape(-1)
        assert np.linalg.norm(r) / np.linalg.norm(b) <= res


@pytest.mark.gpu
def test_gpu_solvers_follow_the_same_rules(system):
    from gpu_util import Dev, init_lattice
    import openstaple_amd as ost
    from openstaple_amd import lib

    o, u, ph, K, b = system
    m = 0.1
    A = m * m * np.eye(K.shape[0]) + K
    bv = b.reshape(-1)
    init_lattice((4, 4, 4, 4))
    du, dph, db = Dev(u), Dev(ph), Dev(b)
    scratch = [Dev(nbytes=b.nbytes) for _ in range(4)]
    fp = ost.ferm_param(m, dph.p)

    def gpu_cg(res, max_cg=10000, restart=10000):
        lib.ostp_set_inverter_tricks(0, 0, 0.1, restart)
        x = Dev(nbytes=b.nbytes)
        cg = C.c_int(0)
        st = lib.ker_invert_openacc(du.p, C.byref(fp), x.p, db.p, res, *(s.p for s in scratch), max_cg, 0.0,
                                    C.byref(cg))
        ost.check(0, "ker_invert_openacc")
        return st, cg.value

    hist = cg_history(A, bv, np.zeros_like(bv), ITERS)
    res, first95, first100 = pick_res(hist)
    assert gpu_cg(res) == (1, first95) and first95 > first100
    R = 40
    hr = cg_history(A, bv, np.zeros_like(bv), ITERS, restart_every=R)
    res_r = pick_res(hr)[0]
    assert gpu_cg(res_r, restart=R) == (1, cg_count(hr, res_r, R, 10000))
    assert gpu_cg(res, max_cg=7)[1] == first95
    assert gpu_cg(1e-14, max_cg=60, restart=R) == (0, 80)
    lib.ostp_set_inverter_tricks(0, 0, 0.1, 10000)

    shifts = [0.0, 0.004, 0.05, 0.6, 4.0]
    sh = shifted_residual_history(A, bv, shifts, ITERS)
    for k in range(int(np.argmax(sh[0] < 1e-8)), ITERS - 5):
        res = sh[0][k] / 0.975
        if separated(sh.ravel(), res * SAFETY) and np.argmax(sh[0] < SAFETY * res) > np.argmax(sh[0] < res):
            break
    n, _ = cgm_count(sh, res)
    ra = ost.RationalApprox()
    ra.approx_order = len(shifts)
    for k, sg in enumerate(shifts):
        ra.RA_b[k] = sg
        ra.RA_a[k] = 1.0
    order = len(shifts)
    xs, sf = Dev(nbytes=order * b.nbytes), Dev(nbytes=order * b.nbytes)
    cg = C.c_int(0)
    st = lib.multishift_invert(du.p, C.byref(fp), C.byref(ra), xs.p, db.p, res, *(s.p for s in scratch), sf.p,
                               10000, C.byref(cg))
    ost.check(0, "multishift_invert")
    assert (st, cg.value) == (1, n)
