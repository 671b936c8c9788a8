"""The Remez generator (openstaple_amd/remez.py) against the approximation
files the reference's own GMP Remez wrote (saved_approxs/ and the tolerance-
driven files in tests/golden), and the committed x^(1/4), x^(-1/2) fixtures
against the minimax property itself.

Bar:
  * fixed order (rationalapprox_calc_fixed_order.cc): x^(-1/4) of orders 9
    and 19 on [4e-7, 1]: Error, a0, every residue and pole within 1e-8
    relative of the reference's file;
  * by tolerance (rationalapprox_calc.cc): the order the reference's guess
    ln(tol) ln(lambda_min)/19 + 2 and its increments arrive at, and the
    coefficients, for three files at lambda_min = 4.673618e-4;
  * the committed order-19 fixtures of x^(1/4) and x^(-1/2) on [1e-7, 1]: the
    relative error of the partial fractions, evaluated at 40 digits on a
    logarithmic grid, equioscillates: exactly 2n + 2 = 40 alternating extrema,
    all of magnitude Error within 1e-3 (the characterisation of the unique
    minimax approximation), with no larger error anywhere;
  * a file written by the generator reads back through the library's
    rationalapprox_read_custom_nomefile into the same struct.
"""
import ctypes as C
import pathlib

import mpmath as mp
import numpy as np
import pytest

import openstaple_amd as ost
from openstaple_amd import lib, remez

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def coeffs(ra):
    n = ra.approx_order
    return np.array([ra.RA_a0] + list(ra.RA_a[:n])), np.array(list(ra.RA_b[:n]))


def assert_same(ra, ref, tol=1e-8):
    assert ra.approx_order == ref.approx_order
    assert (ra.exponent_num, ra.exponent_den) == (ref.exponent_num, ref.exponent_den)
    assert ra.error == pytest.approx(ref.error, rel=tol)
    a, b = coeffs(ra)
    ra_, rb_ = coeffs(ref)
    assert np.all(np.abs(a - ra_) <= tol * np.abs(ra_)), np.max(np.abs(a - ra_) / np.abs(ra_))
    assert np.all(np.abs(b - rb_) <= tol * np.abs(rb_)), np.max(np.abs(b - rb_) / np.abs(rb_))


@pytest.mark.parametrize("order,fname", [(9, "approx_-1_over_4_order_9_mloglm_6.4.REMEZ"),
                                         (19, "approx_-1_over_4_order_19_mloglm_6.4.REMEZ")])
def test_fixed_order_matches_reference_files(order, fname):
    ref = ost.read_remez(str(GOLDEN / fname))
    ra = remez.generate(-1, 4, ref.lambda_min, order)
    assert_same(ra, ref)


@pytest.mark.parametrize("fname,num,den,mlogerr", [
    ("approx_-1_over_4_mlogerr_3.8_mloglm_3.3.REMEZ", -1, 4, 3.8),
    ("approx_-1_over_4_mlogerr_7.8_mloglm_3.3.REMEZ", -1, 4, 7.8),
    ("approx_1_over_8_mlogerr_8.1_mloglm_3.3.REMEZ", 1, 8, 8.1),
])
def test_tolerance_mode_matches_reference_files(fname, num, den, mlogerr):
    ref = ost.read_remez(str(GOLDEN / fname))
    tol = 10.0 ** (-mlogerr)
    ra = remez.generate_to_tolerance(tol, num, den, ref.lambda_min)
    assert ra.error <= tol
    assert_same(ra, ref)


def relerr_extrema(ra, npts=20000):
    n = ra.approx_order
    p = mp.mpf(ra.exponent_num) / ra.exponent_den
    with mp.workdps(40):
        la, lb = mp.log(mp.mpf(ra.lambda_min)), mp.log(mp.mpf(ra.lambda_max))
        xs = [mp.exp(la + (lb - la) * i / (npts - 1)) for i in range(npts)]
        e = []
        for x in xs:
            r = mp.mpf(ra.RA_a0) + mp.fsum(mp.mpf(ra.RA_a[k]) / (x + mp.mpf(ra.RA_b[k])) for k in range(n))
            e.append(float(r / x ** p - 1))
    e = np.array(e)
    # sign runs and the extreme value of each
    runs, cur = [], [np.sign(e[0]), abs(e[0])]
    for v in e[1:]:
        if np.sign(v) == cur[0]:
            cur[1] = max(cur[1], abs(v))
        else:
            runs.append(cur)
            cur = [np.sign(v), abs(v)]
    runs.append(cur)
    return e, runs


@pytest.mark.parametrize("fname", ["approx_1_over_4_order_19_mloglm_7.0.REMEZ",
                                   "approx_-1_over_2_order_19_mloglm_7.0.REMEZ"])
def test_committed_fixtures_equioscillate(fname):
    ra = ost.read_remez(str(GOLDEN / fname))
    assert ra.approx_order == 19 and ra.lambda_min == pytest.approx(1e-7) and ra.lambda_max == 1.0
    e, runs = relerr_extrema(ra)
    assert len(runs) == 2 * ra.approx_order + 2
    peaks = np.array([m for _, m in runs])
    # the inverse of an approximation of x^|p| keeps its extrema to first order in E
    assert np.all(np.abs(peaks - ra.error) <= 1e-3 * ra.error), (peaks.min(), peaks.max(), ra.error)
    assert np.max(np.abs(e)) <= ra.error * (1 + 1e-3)
    assert np.all(np.array(ra.RA_b[:19]) > 0) and np.all(np.diff(ra.RA_b[:19]) > 0)


def test_written_file_reads_back(tmp_path):
    ra = remez.generate(1, 4, 1e-3, 6)
    path = tmp_path / "approx_test.REMEZ"
    remez.write(str(path), ra)
    back = ost.RationalApprox()
    assert lib.rationalapprox_read_custom_nomefile(C.byref(back), str(path).encode()) == 0
    assert bytes(back) == bytes(ra)
