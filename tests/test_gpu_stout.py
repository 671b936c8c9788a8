"""GPU parity of stout smearing, its force chain and unitarisation against the
oracle (which tests/test_oracle_stout.py pins to the reference's own
stouting.c).  Bar: FP64, 1e-13 relative on links, Lambda and Sigma."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import openstaple_amd as ost  # noqa: E402
from openstaple_amd import lib  # noqa: E402
from gpu_util import Dev, init_lattice  # noqa: E402
from oracle import Oracle, ta_view  # noqa: E402

RHO = 0.15
SHAPES = [(4, 4, 4, 4), (8, 8, 8, 8), (4, 6, 8, 10)]


def maxrel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(autouse=True)
def stout_params():
    lib.ostp_set_stout(RHO, 2, 0.0, 0, 0, 1)
    yield


@pytest.mark.parametrize("shape", SHAPES)
def test_stout_isotropic_and_parking_outputs(shape):
    o = Oracle(*shape)
    init_lattice(shape)
    u = o.gen_cold(0.7, 3)
    du = Dev(u)
    dout, dst, dex = Dev(nbytes=u.nbytes), Dev(nbytes=u.nbytes), Dev(nbytes=u.nbytes)
    qa0 = o.ta_zero()
    dqa = Dev(nbytes=qa0.nbytes)
    lib.stout_isotropic(du.p, dout.p, dst.p, dex.p, dqa.p, 0)
    ost.check(0, "stout_isotropic")
    ref = o.stout(u, RHO)
    got = dout.get(u)
    assert maxrel(o.links(got)[:, :2], o.links(ref)[:, :2]) < 1e-13
    st = o.staples(u)
    assert maxrel(o.links(dst.get(u)), o.links(st)) < 1e-13
    q = o.qa(u, st, RHO)
    c, d = ta_view(o, dqa.get(q))
    rc, rd = ta_view(o, q)
    assert maxrel(c, rc) < 1e-12 and maxrel(d, rd) < 1e-12


def test_stout_wrapper_two_levels():
    shape = (8, 8, 8, 8)
    o = Oracle(*shape)
    init_lattice(shape)
    u = o.gen_cold(0.5, 1)
    du = Dev(u)
    arr = Dev(nbytes=2 * u.nbytes)
    lib.stout_wrapper(du.p, arr.p, 0)
    ost.check(0, "stout_wrapper")
    both = arr.get(np.zeros(2 * u.nbytes, dtype=np.uint8))
    l1, l2 = both[: u.nbytes], both[u.nbytes:]
    r1 = o.stout(u, RHO)
    r2 = o.stout(r1, RHO)
    assert maxrel(o.links(l1)[:, :2], o.links(r1)[:, :2]) < 1e-13
    assert maxrel(o.links(l2)[:, :2], o.links(r2)[:, :2]) < 1e-13


@pytest.mark.parametrize("shape", SHAPES)
def test_sigma_chain(shape):
    o = Oracle(*shape)
    init_lattice(shape)
    rng = np.random.default_rng(5)
    u = o.gen_cold(0.7, 9)
    sp = o.gauge()
    S = o.links(sp)
    S[:] = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    ref_s, ref_l, ref_q = o.sigma_prime_to_sigma(sp, u, RHO)
    du, ds = Dev(u), Dev(sp)
    dl, dq, dt = Dev(nbytes=ref_l.nbytes), Dev(nbytes=ref_q.nbytes), Dev(nbytes=u.nbytes)
    lib.compute_sigma_from_sigma_prime_backinto_sigma_prime(ds.p, dl.p, dq.p, du.p, dt.p, 0)
    ost.check(0, "sigma chain")
    c, d = ta_view(o, dl.get(ref_l))
    rc, rd = ta_view(o, ref_l)
    assert maxrel(c, rc) < 1e-12 and maxrel(d, rd) < 1e-12
    assert maxrel(o.links(ds.get(sp)), o.links(ref_s)) < 1e-12
    # the reference's two-step API: compute_lambda with the given QA, then compute_sigma
    ds.put(sp)
    lib.compute_lambda(dl.p, ds.p, du.p, dq.p, dt.p)
    lib.compute_sigma(dl.p, du.p, ds.p, dq.p, dt.p, 0)
    ost.check(0, "lambda/sigma")
    assert maxrel(o.links(ds.get(sp)), o.links(ref_s)) < 1e-12


def test_unitarize():
    shape = (8, 8, 8, 8)
    o = Oracle(*shape)
    init_lattice(shape)
    rng = np.random.default_rng(2)
    u = o.gen_cold(0.5, 4)
    L = o.links(u)
    L[:] += 0.03 * (rng.standard_normal(L.shape) + 1j * rng.standard_normal(L.shape))
    du = Dev(u)
    lib.unitarize_conf(du.p)
    ost.check(0, "unitarize")
    ref = o.unitarize(u)
    assert maxrel(o.links(du.get(u))[:, :2], o.links(ref)[:, :2]) < 1e-15


def test_stout_full_size_properties():
    """32^4 smearing of the benchmark field: level 1 equals the oracle on every
    link, and the level-2 links stay SU(3) (rows 0,1 orthonormal)."""
    shape = (32, 32, 32, 32)
    L = init_lattice(shape)
    lat = ost.Lattice(*shape)
    gb = 8 * lat.su3_bytes
    u = Dev(nbytes=gb)
    lib.ostp_generate_Conf_cold(u.p, 0.5, 1)
    arr = Dev(nbytes=2 * gb)
    lib.stout_wrapper(u.p, arr.p, 0)
    ost.check(0, "stout 32^4")
    o = Oracle(*shape)
    host = arr.get(np.zeros(2 * gb, dtype=np.uint8))
    lev1 = o.links(host[:gb].copy())[:, :2]
    lev2 = o.links(host[gb:].copy())[:, :2]
    ref1 = o.links(o.stout(o.gen_cold(0.5, 1), RHO))[:, :2]
    d1 = np.abs(lev1 - ref1).max(axis=(1, 2))
    assert d1.max() < 1e-13, (d1.max(), np.unravel_index(np.argmax(d1), d1.shape))
    rows = np.transpose(lev2, (0, 3, 1, 2)).reshape(-1, 2, 3)
    gram = rows @ np.conj(np.transpose(rows, (0, 2, 1)))
    dev = np.abs(gram - np.eye(2)).max(axis=(1, 2))
    worst = int(np.argmax(dev))
    assert dev[worst] < 1e-12, (dev[worst], divmod(worst, o.sizeh), rows[worst])
