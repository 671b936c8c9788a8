"""ker_invert_openacc_noalloc through the C ABI: the same solutions (bit for
bit) and iteration counts as ker_invert_openacc, on two calls in a row, the
second on a smaller lattice; larger lattices than its static scratch holds
are rejected with an error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import openstaple_amd as ost  # noqa: E402
from openstaple_amd import lib  # noqa: E402
from gpu_util import Dev, init_lattice  # noqa: E402
from oracle import Oracle  # noqa: E402


def solve(fn, shape, seed, mass, res):
    o = Oracle(*shape)
    init_lattice(shape)
    lib.ostp_set_inverter_tricks(0, 0, 0.1, 10000)
    u, ph, b = o.gen_cold(0.5, seed), o.phases(), o.gen_gauss(seed + 1)
    du, dph, db = Dev(u), Dev(ph), Dev(b)
    x = Dev(nbytes=b.nbytes)
    scratch = [Dev(nbytes=b.nbytes) for _ in range(4)]
    fp = ost.ferm_param(mass, dph.p)
    cg = C.c_int(0)
    st = fn(du.p, C.byref(fp), x.p, db.p, res, *(s.p for s in scratch), 10000, 0.0, C.byref(cg))
    ost.check(0, "solve")
    r = x.get(b)
    rs = scratch[0].get(b)  # loc_r: the last recursive residual
    return st, cg.value, r, rs


def test_noalloc_cg_matches_ker_invert_openacc():
    for shape, seed, mass in (((8, 8, 8, 8), 3, 0.05), ((4, 4, 4, 8), 5, 0.1)):  # second call: smaller lattice
        ref = solve(lib.ker_invert_openacc, shape, seed, mass, 1e-10)
        got = solve(lib.ker_invert_openacc_noalloc, shape, seed, mass, 1e-10)
        assert ref[0] == got[0] == 1
        assert ref[1] == got[1] > 10
        assert np.array_equal(ref[2], got[2])
        assert np.array_equal(ref[3], got[3])


def test_noalloc_rejects_lattices_beyond_its_scratch():
    init_lattice((64, 64, 64, 66))  # 2^23 + 2^18 half-sites (no fields allocated; errors recorded, not fatal)
    fp = ost.FermParam()
    cg = C.c_int(-1)
    st = lib.ker_invert_openacc_noalloc(None, C.byref(fp), None, None, 1e-10, None, None, None, None, 100, 0.0,
                                        C.byref(cg))
    msg = lib.ostp_last_error().decode()
    lib.ostp_clear_error()
    assert st == 0 and "half-sites" in msg and cg.value == -1
