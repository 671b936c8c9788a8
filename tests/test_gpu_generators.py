"""GPU field generators vs the oracle's restatement of the same Philox rules:
the integer stream and the Z2 noise are bit-exact; derived doubles agree to
rounding (device vs host libm); links are SU(3) to machine precision."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import openstaple_amd as ost  # noqa: E402
from openstaple_amd import lib  # noqa: E402
from gpu_util import Dev, init_lattice, rel  # noqa: E402
from oracle import Oracle  # noqa: E402

SHAPE = (8, 8, 8, 8)


@pytest.fixture(scope="module")
def lat():
    return Oracle(*SHAPE), init_lattice(SHAPE)


def test_z2_bitexact(lat):
    o, _ = lat
    ref = o.gen_z2(77, 3)
    d = Dev(nbytes=ref.nbytes)
    lib.ostp_generate_vec3_soa_z2noise(d.p, 77, 3)
    ost.check(0, "z2")
    np.testing.assert_array_equal(d.get(ref), ref)


def test_gauss(lat):
    o, _ = lat
    ref = o.gen_gauss(1234, 0)
    d = Dev(nbytes=ref.nbytes)
    lib.ostp_generate_vec3_soa_gauss(d.p, 1234, 0)
    ost.check(0, "gauss")
    got = d.get(ref)
    assert np.max(np.abs(got - ref)) < 1e-14
    assert abs(np.mean(np.abs(got) ** 2) - 1.0) < 0.02


@pytest.mark.parametrize("kind", ["cold", "expH"])
def test_gauge_generators(lat, kind):
    o, _ = lat
    ref = o.gen_cold(0.5, 99) if kind == "cold" else o.gen_expH(0.3, 1)
    d = Dev(nbytes=ref.nbytes)
    if kind == "cold":
        lib.ostp_generate_Conf_cold(d.p, 0.5, 99)
    else:
        lib.ostp_generate_Conf_expH(d.p, 0.3, 1)
    ost.check(0, kind)
    got = d.get(ref)
    G = o.links(got)
    R = o.links(ref)
    assert np.max(np.abs(G - R)) < 1e-13
    m = np.transpose(G, (0, 3, 1, 2)).reshape(-1, 3, 3)
    eye = np.eye(3)
    assert np.max(np.abs(m @ np.conj(np.transpose(m, (0, 2, 1))) - eye)) < 1e-13
    assert np.max(np.abs(np.linalg.det(m) - 1.0)) < 1e-13
