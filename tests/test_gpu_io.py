"""Field files from and to device arrays (ostp_save_conf, ostp_read_conf):
the library's su3_soa[8] (with or without the PAR_TEMP K plane) staged to
the reference's global layout.

Bar: the ILDG file written from a device field is byte-identical to the one
print_su3_soa_ildg_binary writes from the same field in host memory; reading
it (or the ASCII file) back into device memory restores rows 0 and 1 (bit
for bit for ILDG, to the printed decimals for ASCII) and the rebuilt row 2,
leaves the K plane alone, and updates a registered host array's host copy.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import openstaple_amd as ost  # noqa: E402
from openstaple_amd import lib  # noqa: E402
from gpu_util import Dev, init_lattice  # noqa: E402
from oracle import Oracle  # noqa: E402


def planes(u, o, pt):
    """(8, 9, sizeh) complex colour planes of an su3_soa[8] byte buffer"""
    per = (144 + (8 if pt else 0)) * o.sizeh
    return np.stack([np.frombuffer(u[l * per:l * per + 144 * o.sizeh].tobytes(), dtype=np.complex128)
                     .reshape(9, o.sizeh) for l in range(8)])


@pytest.mark.parametrize("par_temp,use_ildg", [(0, 1), (1, 1), (1, 0)])
def test_save_and_read_conf_on_device(tmp_path, par_temp, use_ildg):
    shape = (4, 4, 4, 6)
    o = Oracle(*shape, par_temp=bool(par_temp))
    u = o.gen_expH(0.4, 3)
    if par_temp:
        o.init_k(u, 0.6, 2, (1, 2, 0))
    init_lattice(shape, par_temp=bool(par_temp))
    du = Dev(u)
    path = str(tmp_path / ("c.ildg" if use_ildg else "c.txt")).encode()
    assert lib.ostp_save_conf(du.p, path, 9, use_ildg) == 0
    P = planes(u, o, par_temp)
    if use_ildg:
        host = np.ascontiguousarray(P)  # the global layout: 9 planes per link array, no K
        path2 = str(tmp_path / "h.ildg").encode()
        assert lib.print_su3_soa_ildg_binary(host.ctypes.data, path2, 9) == 0
        assert open(path, "rb").read() == open(path2, "rb").read()
    d2 = Dev(nbytes=u.nbytes)
    cid = C.c_int(0)
    assert lib.ostp_read_conf(d2.p, path, C.byref(cid), use_ildg) == 0
    assert cid.value == 9
    back = d2.get(u)
    B = planes(back, o, par_temp)
    if use_ildg:
        assert np.array_equal(B[:, :6], P[:, :6])
    else:
        assert np.max(np.abs(B[:, :6] - P[:, :6])) < 1e-17
    assert np.max(np.abs(B[:, 6:] - P[:, 6:])) < 1e-15
    if par_temp:
        per = 152 * o.sizeh
        k = np.stack([np.frombuffer(back[l * per + 144 * o.sizeh:(l + 1) * per].tobytes(), dtype=np.float64)
                      for l in range(8)])
        assert np.all(k == 0.0)
    # a registered host array: device mirror and host copy both receive the field
    h = np.zeros_like(u)
    pr = ost.Present(h, copyin=True)
    assert lib.ostp_read_conf(pr.p, path, C.byref(cid), use_ildg) == 0
    assert np.array_equal(planes(h, o, par_temp)[:, :6], B[:, :6])
    dev_copy = np.zeros_like(u)
    pr.update_self()
    assert np.array_equal(planes(pr.a, o, par_temp), planes(h, o, par_temp))
    pr.delete()
