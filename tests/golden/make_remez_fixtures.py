"""Writes the order-19 rational approximations of x^(1/4) and x^(-1/2) on
[1e-7, 1] that the 2+1-flavour RHMC at m_l = 0.0015 needs (heat bath and
MD / final action of the light quark, degeneracy 2, one pseudofermion: the
exponents +deg/8 and -deg/4 of Include/fermion_parameters.c:66-72).  The
reference's saved approximations stop at lambda_min = 4e-7, which does not
cover m_l^2 = 2.25e-6 once rescaled to lambda_max ~ 17.

    python tests/golden/make_remez_fixtures.py   (about a minute)
"""
import pathlib
import sys

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from openstaple_amd import remez  # noqa: E402

FIXTURES = [(1, 4, 1e-7, 19), (-1, 2, 1e-7, 19)]


def name(num, den, lmin, order):
    return "approx_%d_over_%d_order_%d_mloglm_7.0.REMEZ" % (num, den, order)


if __name__ == "__main__":
    for num, den, lmin, order in FIXTURES:
        ra = remez.generate(num, den, lmin, order)
        remez.write(str(HERE / name(num, den, lmin, order)), ra)
        print(name(num, den, lmin, order), "error %.6e" % ra.error)
