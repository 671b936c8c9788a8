"""The library's host API for rational approximations (rationalapprox.c) against
the reference's own rationalapprox.c compiled in oracle/_ref: every .REMEZ
fixture read, rescaled, renormalised, evaluated and written bit for bit alike,
and the file names alike.  CPU only (host code; no device work)."""
import ctypes as C

import numpy as np
import pytest

from conftest import ref_lib
from oracle import RatApprox
import pathlib

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
REMEZ = sorted(GOLDEN.glob("*.REMEZ"))


def _headerless(path):
    return "Order:" not in path.read_text()


@pytest.fixture(scope="module")
def libs():
    ref = ref_lib((4, 4, 4, 4))
    if ref is None:
        pytest.skip("oracle/_ref not built (make -C oracle ref)")
    import openstaple_amd as ost
    ost.lib.ostp_set_abort_on_error(0)
    for L in (ref,):
        L.rationalapprox_read_custom_nomefile.argtypes = [C.c_void_p, C.c_char_p]
        L.rational_approx_evaluate.restype = C.c_double
        L.rational_approx_evaluate.argtypes = [C.c_void_p, C.c_double]
        L.rational_approx_filename.restype = C.c_void_p
        L.rational_approx_filename.argtypes = [C.c_double, C.c_int, C.c_int, C.c_double]
        L.rational_approx_filename_old.restype = C.c_void_p
        L.rational_approx_filename_old.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
        L.rationalapprox_save.argtypes = [C.c_char_p, C.c_void_p]
        L.rescale_rational_approximation.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.renormalize_rational_approximation.argtypes = [C.c_void_p, C.c_void_p]
    yield ost.lib, ref
    ost.lib.ostp_set_abort_on_error(1)


def _bytes(ra):
    return bytes(memoryview(ra))


def _read(L, path):
    ra = RatApprox()
    rc = L.rationalapprox_read_custom_nomefile(C.byref(ra), str(path).encode())
    return rc, ra


@pytest.mark.parametrize("path", [p for p in REMEZ if not _headerless(p)], ids=lambda p: p.name)
def test_read_rescale_evaluate_save_bitexact(libs, path, tmp_path):
    lib, ref = libs
    rc_l, ra_l = _read(lib, path)
    rc_r, ra_r = _read(ref, path)
    assert rc_l == rc_r == 0
    assert _bytes(ra_l) == _bytes(ra_r)
    # rescale to a range the approximation covers: min = 2 x the rescaled lambda_min
    vmax = 5.7
    vmin = 2.0 * ra_l.lambda_min * vmax * 1.05
    mm = (C.c_double * 2)(vmin, vmax)
    out_l, out_r = RatApprox(), RatApprox()
    lib.rescale_rational_approximation(C.byref(ra_l), C.byref(out_l), mm)
    ref.rescale_rational_approximation(C.byref(ra_r), C.byref(out_r), mm)
    assert _bytes(out_l) == _bytes(out_r)
    ren_l, ren_r = RatApprox(), RatApprox()
    lib.renormalize_rational_approximation(C.byref(out_l), C.byref(ren_l))
    ref.renormalize_rational_approximation(C.byref(out_r), C.byref(ren_r))
    assert _bytes(ren_l) == _bytes(ren_r)
    for x in np.geomspace(out_l.lambda_min, out_l.lambda_max, 13):
        assert lib.rational_approx_evaluate(C.byref(out_l), x) == ref.rational_approx_evaluate(C.byref(out_r), x)
    fl, fr = tmp_path / "lib.REMEZ", tmp_path / "ref.REMEZ"
    lib.rationalapprox_save(str(fl).encode(), C.byref(out_l))
    ref.rationalapprox_save(str(fr).encode(), C.byref(out_r))
    assert fl.read_bytes() == fr.read_bytes()
    # and the written file reads back to the same struct
    rc, back = _read(lib, fl)
    assert rc == 0 and back.approx_order == out_l.approx_order


def test_file_names_alike(libs):
    lib, ref = libs
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]

    def take(p):
        s = C.string_at(p).decode()
        libc.free(p)
        return s

    for err, num, den, lmin in [(1e-8, -1, 4, 4e-7), (3e-4, 1, 8, 5e-4), (2.5e-10, -1, 2, 1e-6)]:
        assert take(lib.rational_approx_filename(err, num, den, lmin)) == \
            take(ref.rational_approx_filename(err, num, den, lmin))
    for order, num, den, lmin in [(19, -1, 4, 4e-7), (9, 1, 8, 4e-7)]:
        assert take(lib.rational_approx_filename_old(order, num, den, lmin)) == \
            take(ref.rational_approx_filename_old(order, num, den, lmin))


def test_missing_and_malformed_files(libs, tmp_path):
    """A missing file returns 1 (as the reference); a file without the header
    is an error (the reference exits; the library reports it and returns 2)."""
    lib, _ = libs
    rc, _ = _read(lib, tmp_path / "nope.REMEZ")
    assert rc == 1
    for path in [p for p in REMEZ if _headerless(p)]:
        rc, _ = _read(lib, path)
        assert rc == 2, path.name
