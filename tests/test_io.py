"""Field files on the host (no GPU): the reference's ASCII gauge and vec3_soa
formats and the ILDG binary format.

Bar:
  * ASCII conf (io.c:35-207): the header line and the "(re, im)  " rows
    carry the printf formats of the reference, row 2 rebuilt from rows 0, 1;
    write -> read returns the field to the 18 printed decimals; matrices
    written with a staggered sign (det = -1) are negated back on reading;
  * ILDG (io.c:257-547): an independent reader written from the LIME / ILDG
    specification (144-byte big-endian record headers, magic 0x456789ab,
    8-byte padding; ildg-binary-data with t slowest, x fastest, directions
    x, y, z, t, 3x3 row-major complex, big-endian FP64) finds the records
    ildg-format (with lx..lt), MD_traj, input-file, ildg-binary-data,
    ildg-data-lfn and every matrix of the field at its place; the library's
    reader returns rows 0 and 1 bit for bit and row 2 as the rebuilt row;
  * vec3_soa ASCII (dbgtools.c:1193-1236): "%.18lf\\t%.18lf" lines, c0, c1,
    c2 of each site, and the round trip;
  * errors: a missing file returns 1 (conf_id -1), a malformed or mismatched
    one 2.
"""
import ctypes as C
import struct

import numpy as np
import pytest

import openstaple_amd as ost
from openstaple_amd import lib

SHAPE = (4, 2, 4, 6)


@pytest.fixture(autouse=True)
def host_lattice():
    lib.ostp_set_abort_on_error(0)
    assert lib.ostp_set_lattice_host(*SHAPE, 0) == 0
    lib.ostp_clear_error()
    yield


def sizeh():
    return SHAPE[0] * SHAPE[1] * SHAPE[2] * SHAPE[3] // 2


def random_su3(n, rng):
    z = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    q, r = np.linalg.qr(z)
    q = q * (np.diagonal(r, axis1=1, axis2=2) / np.abs(np.diagonal(r, axis1=1, axis2=2)))[:, None, :]
    return q / np.linalg.det(q)[:, None, None] ** (1 / 3)


def field(rng):
    """global_su3_soa[8] as (8, 3, 3, sizeh) complex"""
    sh = sizeh()
    m = random_su3(8 * sh, rng).reshape(8, sh, 3, 3)
    return np.ascontiguousarray(np.transpose(m, (0, 2, 3, 1)))


def parity_index(x, y, z, t):
    return (x + y + z + t) % 2, (x + SHAPE[0] * (y + SHAPE[1] * (z + SHAPE[2] * t))) // 2


def test_ascii_conf_format_and_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    u = field(rng)
    ost.ActionParam.in_dll(lib, "act_params").beta = 5.3
    ost.AllocSettings.in_dll(lib, "alloc_info").NDiffFlavs = 0
    path = str(tmp_path / "conf.txt").encode()
    assert lib.print_su3_soa_ASCII(u.ctypes.data, path, 17, None) == 0
    lines = open(path).read().split("\n")
    assert lines[0] == "%d %d %d %d %f %f %d %d" % (*SHAPE, 5.3, 0.0, 0, 17)
    m = u[0, :, :, 0].copy()
    m[2] = np.conj(np.cross(m[0], m[1]))
    for r in range(2):
        assert lines[1 + r] == "".join("(%.18f, %.18f)  " % (m[r, c].real, m[r, c].imag) for c in range(3))
    # row 2 is the rebuilt row (rounding of the rebuild aside)
    nums = [float(t) for t in lines[3].replace("(", " ").replace(")", " ").replace(",", " ").split()]
    assert np.allclose(np.array(nums[0::2]) + 1j * np.array(nums[1::2]), m[2], rtol=0, atol=1e-15)
    assert lines[3].count("(") == 3 and lines[3].endswith(")  ")
    assert lines[4] == ""
    back = np.zeros_like(u)
    cid = C.c_int(0)
    assert lib.read_su3_soa_ASCII(back.ctypes.data, path, C.byref(cid)) == 0
    assert cid.value == 17
    assert np.max(np.abs(back[:, :2] - u[:, :2])) < 1e-17
    assert np.max(np.abs(back[:, 2] - u[:, 2])) < 1e-15  # the rebuilt row was written


def test_ascii_conf_staggered_sign_is_undone(tmp_path):
    rng = np.random.default_rng(2)
    u = field(rng)
    sh = sizeh()
    vol = 2 * sh
    ph = np.zeros((8, sh))
    flat = ph.reshape(-1)
    flat[rng.permutation(8 * sh)[:vol * 3 // 2]] = 1.0
    path = str(tmp_path / "conf_ph.txt").encode()
    assert lib.print_su3_soa_ASCII(u.ctypes.data, path, 3, ph.ctypes.data) == 0
    back = np.zeros_like(u)
    cid = C.c_int(0)
    assert lib.read_su3_soa_ASCII(back.ctypes.data, path, C.byref(cid)) == 0
    assert np.max(np.abs(back[:, :2] - u[:, :2])) < 1e-17
    assert np.max(np.abs(back[:, 2] - u[:, 2])) < 1e-15
    # a count of signs that is not 3/2 of the volume is reported
    flat[:] = 0
    flat[:5] = 1
    assert lib.print_su3_soa_ASCII(u.ctypes.data, path, 3, ph.ctypes.data) == 1


def lime_records(path):
    data = open(path, "rb").read()
    pos, out = 0, []
    while pos < len(data):
        magic, version, flags, length, typ = struct.unpack(">IHHQ128s", data[pos:pos + 144])
        assert magic == 0x456789AB and version == 1
        name = typ.split(b"\0")[0].decode()
        out.append((name, data[pos + 144:pos + 144 + length]))
        pos += 144 + length + (-length % 8)
    return out


def test_ildg_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    u = field(rng)
    path = str(tmp_path / "conf.ildg").encode()
    inp = (C.c_char * (300 * 500)).in_dll(lib, "input_file_str")
    inp.value = b"Beta 5.3\nNmdSteps 10\n"
    assert lib.print_su3_soa_ildg_binary(u.ctypes.data, path, 42) == 0
    recs = lime_records(path)
    assert [r[0] for r in recs] == ["ildg-format", "MD_traj", "input-file", "ildg-binary-data", "ildg-data-lfn"]
    xml = recs[0][1].decode().rstrip("\0")
    for tag, n in zip(("lx", "ly", "lz", "lt"), SHAPE):
        assert "<%s>%d</%s>" % (tag, n, tag) in xml
    assert "<precision>64</precision>" in xml and "<field>su3gauge</field>" in xml
    assert recs[1][1].rstrip(b"\0") == b"42"
    assert recs[2][1].rstrip(b"\0") == b"Beta 5.3\nNmdSteps 10\n"
    assert len(recs[4][1]) == 0
    nx, ny, nz, nt = SHAPE
    binary = np.frombuffer(recs[3][1], dtype=">f8").astype(np.float64)
    assert binary.size == nx * ny * nz * nt * 4 * 18
    mats = (binary[0::2] + 1j * binary[1::2]).reshape(nt, nz, ny, nx, 4, 3, 3)
    for (x, y, z, t) in [(0, 0, 0, 0), (1, 0, 0, 0), (3, 1, 2, 5), (2, 1, 3, 4)]:
        p, i = parity_index(x, y, z, t)
        for d in range(4):
            assert np.allclose(mats[t, z, y, x, d], u[2 * d + p, :, :, i], rtol=0, atol=1e-15)
    back = np.zeros_like(u)
    cid = C.c_int(0)
    assert lib.read_su3_soa_ildg_binary(back.ctypes.data, path, C.byref(cid)) == 0
    assert cid.value == 42
    assert np.array_equal(back[:, :2], u[:, :2])
    assert np.max(np.abs(back[:, 2] - u[:, 2])) < 1e-15


def test_vec3_ascii(tmp_path):
    rng = np.random.default_rng(4)
    sh = sizeh()
    v = rng.normal(size=(3, sh)) + 1j * rng.normal(size=(3, sh))
    v = np.ascontiguousarray(v * 0.3)
    path = str(tmp_path / "fermion.txt").encode()
    lib.print_vec3_soa(v.ctypes.data, path)
    lines = open(path).read().split("\n")
    assert len(lines) == 3 * sh + 1 and lines[-1] == ""
    assert lines[0] == "%.18f\t%.18f" % (v[0, 0].real, v[0, 0].imag)
    assert lines[4] == "%.18f\t%.18f" % (v[1, 1].real, v[1, 1].imag)
    back = np.zeros_like(v)
    assert lib.read_vec3_soa(back.ctypes.data, path) == 0
    assert np.max(np.abs(back - v)) < 1e-17


def test_missing_and_malformed_files(tmp_path):
    u = np.zeros((8, 3, 3, sizeh()), dtype=np.complex128)
    cid = C.c_int(7)
    assert lib.read_su3_soa_ASCII(u.ctypes.data, str(tmp_path / "none").encode(), C.byref(cid)) == 1
    assert cid.value == -1
    cid.value = 7
    assert lib.read_su3_soa_ildg_binary(u.ctypes.data, str(tmp_path / "none").encode(), C.byref(cid)) == 1
    assert cid.value == -1
    assert lib.read_vec3_soa(u.ctypes.data, str(tmp_path / "none").encode()) == 1
    bad = tmp_path / "short.txt"
    bad.write_text("%d %d %d %d 5.3 0.0 0 1\n(1.0, 0.0) (0.0, 0.0)\n" % SHAPE)
    assert lib.read_su3_soa_ASCII(u.ctypes.data, str(bad).encode(), C.byref(cid)) == 2
    wrong = tmp_path / "wrong.txt"
    wrong.write_text("4 4 4 4 5.3 0.0 0 1\n")
    assert lib.read_su3_soa_ASCII(u.ctypes.data, str(wrong).encode(), C.byref(cid)) == 2
    lib.ostp_clear_error()
    # an ILDG file of another lattice
    rng = np.random.default_rng(5)
    uf = field(rng)
    path = str(tmp_path / "c.ildg").encode()
    assert lib.print_su3_soa_ildg_binary(uf.ctypes.data, path, 1) == 0
    assert lib.ostp_set_lattice_host(4, 2, 4, 4, 0) == 0
    assert lib.read_su3_soa_ildg_binary(u.ctypes.data, path, C.byref(cid)) == 2
    lib.ostp_clear_error()
