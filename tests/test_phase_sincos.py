"""Host check of phase_sincos (openstaple_amd/csrc/ostp_phase.h), the sin/cos
that Dslash and the fermion-force backfield kernels use for the U(1) phases,
against the C library's sin and cos.  The header is plain C, so gcc compiles
the same source the HIP kernels include."""
import pathlib
import subprocess

import pytest

HDR = pathlib.Path(__file__).resolve().parents[1] / "openstaple_amd" / "csrc" / "ostp_phase.h"

PROG = r"""
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include "ostp_phase.h"

static uint64_t st = 0x9E3779B97F4A7C15ull;
static double urand(void) {  /* splitmix64 -> [0,1) */
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

int main(void) {
    const double ranges[] = {0.8, 3.5, 100.0, 1e4, 1e6, 1e8};
    for (int k = 0; k < 6; k++) {
        double me = 0.0;
        for (int i = 0; i < 400000; i++) {
            const double a = ranges[k] * (2.0 * urand() - 1.0);
            double s, c;
            phase_sincos(a, &s, &c);
            const double e = fmax(fabs(s - sin(a)), fabs(c - cos(a)));
            if (e > me) me = e;
        }
        printf("range %g %.6e\n", ranges[k], me);
    }
    /* multiples of pi/2 (staggered phases, antiperiodic boundaries) */
    int mism = 0;
    double me = 0.0;
    for (int k = -4000; k <= 4000; k++) {
        const double a = k * (M_PI / 2);
        double s, c;
        phase_sincos(a, &s, &c);
        if (s != sin(a) || c != cos(a)) mism++;
        me = fmax(me, fmax(fabs(s - sin(a)), fabs(c - cos(a))));
    }
    printf("halfpi %d %.6e\n", mism, me);
    /* exact special values */
    double s, c;
    phase_sincos(0.0, &s, &c);
    printf("zero %d\n", s == 0.0 && c == 1.0);
    phase_sincos(M_PI, &s, &c);
    printf("pi %d\n", s == sin(M_PI) && c == -1.0);
    return 0;
}
"""


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    d = tmp_path_factory.mktemp("phase")
    src = d / "t.c"
    src.write_text(PROG)
    exe = d / "t"
    subprocess.run(["gcc", "-O2", "-std=gnu99", "-ffp-contract=off", f"-I{HDR.parent}", str(src), "-o", str(exe),
                    "-lm"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        k, *v = line.split()
        res[k if k != "range" else f"range {v[0]}"] = v if k != "range" else float(v[1])
    return res


@pytest.mark.parametrize("rng", ["0.8", "3.5", "100", "10000", "1e+06", "1e+08"])
def test_accuracy_against_libm(results, rng):
    # 1.2e-16 absolute: one ulp of a value near 1 is 2.2e-16
    assert results[f"range {rng}"] <= 1.2e-16


def test_multiples_of_half_pi(results):
    # the unit component matches exactly; the near-zero one (sin k*pi ~ 1e-13)
    # may differ from the C library's in its last bits only
    mism, err = results["halfpi"]
    assert float(err) <= 1e-27
    assert results["zero"] == ["1"] and results["pi"] == ["1"]
