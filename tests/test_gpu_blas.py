"""GPU parity of the exported BLAS, reduction and conversion entry points
(fermionic_utilities.c:126-419, float_double_conv.c) through the C ABI.

Pin: the reference's own fermionic_utilities.c, compiled from the reference
tree into oracle/_ref (ctypes, host memory), on the same inputs.  Reductions
within 1e-13 relative (the reference sums in loop order, the GPU in a fixed
tree); element-wise operations within 2 ulp (the GPU may fuse a multiply-add
the x86 build rounds twice).  Conversions: bit-exact against IEEE
round-to-nearest (numpy astype), which is what a C cast does.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import openstaple_amd as ost  # noqa: E402
from openstaple_amd import lib  # noqa: E402
from conftest import ref_lib  # noqa: E402
from gpu_util import Dev, init_lattice  # noqa: E402

SHAPES = [(8, 8, 8, 8), (4, 6, 8, 10)]


def vec(sizeh, seed):
    r = np.random.default_rng(seed)
    return (r.standard_normal(3 * sizeh) + 1j * r.standard_normal(3 * sizeh)).astype(np.complex128)


def close_ulp(got, ref, scale):
    """|got - ref| within 2 ulp of the operands' magnitude, per real entry
    (a fused multiply-add may differ from two roundings by that much)."""
    assert np.all(np.abs(got.view(np.float64) - ref.view(np.float64)) <= 4.5e-16 * scale)


@pytest.fixture(params=SHAPES, ids=lambda s: "x".join(map(str, s)))
def setup(request):
    shape = request.param
    R = ref_lib(shape)
    if R is None:
        pytest.skip("reference build (oracle/_ref) missing")
    R.scal_prod_global.restype = ost.DComplex
    R.scal_prod_global.argtypes = [C.c_void_p, C.c_void_p]
    R.l2norm2_global.restype = C.c_double
    R.l2norm2_global.argtypes = [C.c_void_p]
    R.real_scal_prod_global.restype = C.c_double
    R.real_scal_prod_global.argtypes = [C.c_void_p, C.c_void_p]
    L = init_lattice(shape)
    return shape, R, L.sizeh


def test_reductions(setup):
    shape, R, sizeh = setup
    a, b = vec(sizeh, 1), vec(sizeh, 2)
    da, db = Dev(a), Dev(b)
    ref = R.l2norm2_global(a.ctypes.data)
    got = lib.l2norm2_global(da.p)
    ost.check(0, "l2norm2_global")
    assert abs(got - ref) <= 1e-13 * abs(ref)
    ref = R.real_scal_prod_global(a.ctypes.data, b.ctypes.data)
    got = lib.real_scal_prod_global(da.p, db.p)
    ost.check(0, "real_scal_prod_global")
    assert abs(got - ref) <= 1e-13 * np.sqrt(R.l2norm2_global(a.ctypes.data) * R.l2norm2_global(b.ctypes.data))
    ref = complex(R.scal_prod_global(a.ctypes.data, b.ctypes.data))
    got = complex(lib.scal_prod_global(da.p, db.p))
    ost.check(0, "scal_prod_global")
    norm = np.linalg.norm(a) * np.linalg.norm(b)
    assert abs(got - ref) <= 1e-13 * norm
    # the reference's sum is conj(a) . b: its imaginary part is not symmetric
    assert abs(ref - np.vdot(a, b)) <= 1e-12 * norm and abs(ref.imag) > 1e-3 * norm / np.sqrt(a.size)
    re, im = C.c_double(), C.c_double()
    lib.ostp_scal_prod_global(da.p, db.p, C.byref(re), C.byref(im))
    ost.check(0, "ostp_scal_prod_global")
    assert complex(re.value, im.value) == got


def test_elementwise(setup):
    shape, R, sizeh = setup
    a, b, c = vec(sizeh, 3), vec(sizeh, 4), vec(sizeh, 5)
    fac = C.c_double(-0.73125)
    cases = [
        # (name, argument pattern, buffers written)
        ("combine_in1xfactor_plus_in2", "a f b o"),
        ("combine_add_factor_x_in2_to_in1", "A b f"),
        ("combine_in1xferm_mass_minus_in2", "a f B"),
        ("combine_in1_minus_in2", "a b o"),
        ("assign_in_to_out", "a o"),
        ("set_vec3_soa_to_zero", "A"),
        ("combine_in1_x_fact1_minus_in2_back_into_in2", "a f B"),
        ("combine_in1_minus_in2_allxfact", "a b f o"),
    ]
    for name, pat in cases:
        host = {"a": a.copy(), "b": b.copy(), "o": c.copy()}
        dev = {k: Dev(v) for k, v in host.items()}
        rargs, gargs = [], []
        for t in pat.split():
            if t == "f":
                rargs.append(fac)
                gargs.append(fac)
            else:
                k = t.lower()
                rargs.append(C.c_void_p(host[k].ctypes.data))
                gargs.append(C.c_void_p(dev[k].p))
        getattr(R, name).restype = None
        getattr(R, name)(*rargs)
        getattr(lib, name)(*gargs)
        ost.check(0, name)
        scale = (1 + abs(fac.value)) * np.maximum.reduce([np.abs(v.view(np.float64)) for v in (a, b, c)])
        for k in host:
            got = dev[k].get(host[k])
            close_ulp(got, host[k], scale)
        # the operation changed what it writes (the pins are not vacuous)
        written = [t.lower() for t in pat.split() if t == "o" or t.isupper()]
        for k in written:
            assert not np.array_equal(host[k], {"a": a, "b": b, "o": c}[k]), name


def test_conversions_bitexact(setup):
    shape, R, sizeh = setup
    a = vec(sizeh, 6) * 1e3
    da = Dev(a)
    df = Dev(nbytes=a.nbytes // 2)
    lib.convert_double_to_float_vec3_soa(da.p, df.p)
    ost.check(0, "convert_double_to_float_vec3_soa")
    exp = a.astype(np.complex64)
    got = df.get(exp)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    back = Dev(nbytes=a.nbytes)
    lib.convert_float_to_double_vec3_soa(df.p, back.p)
    ost.check(0, "convert_float_to_double_vec3_soa")
    assert np.array_equal(back.get(a), exp.astype(np.complex128))


@pytest.mark.parametrize("par_temp", [False, True])
@pytest.mark.parametrize("shape", [(4, 6, 8, 10), (2, 3, 3, 3)], ids=lambda s: "x".join(map(str, s)))
def test_gauge_conversion_bitexact(shape, par_temp):
    """All 8 link arrays (9 planes each, and the K plane under parallel
    tempering), on a lattice whose half-volume is odd (2x3x3x3: 27)."""
    L = init_lattice(shape, par_temp=par_temp)
    sizeh = L.sizeh
    per = 9 * sizeh * 16 + (8 * sizeh if par_temp else 0)
    r = np.random.default_rng(7)
    u = r.standard_normal(8 * per // 8)
    du = Dev(u)
    perf = per // 2
    uf = Dev(nbytes=8 * perf)
    lib.ostp_convert_double_to_float_gauge(du.p, uf.p)
    ost.check(0, "ostp_convert_double_to_float_gauge")
    got = uf.get(np.zeros(8 * perf // 4, np.float32))
    exp = u.astype(np.float32)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
