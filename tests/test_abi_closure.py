"""A reference host that drops the files the library replaces still links.

tests/golden/abi_closure.json lists, for every non-static function and global
variable defined in a reference file the library replaces (fermion_matrix.c,
fermionic_utilities.c, io.c, ...), the kept reference files that still call
or reference it (bodies and static inline header functions; generated
single-precision twins included).  Every such symbol must be exported by the
library, or the reference program fails to link.  The fixture is regenerated
from the reference tree by tests/golden/make_abi_closure.py.
"""
import ctypes
import json
import pathlib
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
LIB = ROOT / "openstaple_amd" / "libopenstaple_amd.so"
FIXTURE = ROOT / "tests" / "golden" / "abi_closure.json"
REF = pathlib.Path("/root/reference/src")


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if l.strip()}


def test_every_symbol_the_kept_reference_code_needs_is_exported():
    if not LIB.exists():
        pytest.skip("library not built")
    need = json.loads(FIXTURE.read_text())["needed"]
    assert len(need) >= 20  # the fixture is not empty by accident
    missing = {k: v for k, v in need.items() if k not in exported()}
    assert not missing, missing


def test_integration_guide_drops_the_same_files():
    fx = json.loads(FIXTURE.read_text())
    guide = (ROOT / "INTEGRATION.md").read_text()
    for f in fx["dropped_files"] + fx["dropped_twins_of_kept_files"]:
        assert f in guide, f
    for f, names in fx["dropped_definitions"].items():
        for n in names:
            assert n in guide, (f, n)


@pytest.mark.skipif(not REF.is_dir(), reason="reference tree not present")
def test_fixture_matches_the_reference_tree(tmp_path, monkeypatch):
    sys.path.insert(0, str(ROOT / "tests" / "golden"))
    import make_abi_closure as m

    monkeypatch.setattr(m, "OUT", tmp_path / "abi_closure.json")
    assert m.main() == 0
    new = json.loads((tmp_path / "abi_closure.json").read_text())
    old = json.loads(FIXTURE.read_text())
    assert new["needed"] == old["needed"]
    assert new["removed_definitions"] == old["removed_definitions"]


def test_scanner_finds_definitions_calls_and_pointer_tables(tmp_path, monkeypatch):
    """The scanner on a small hand-written C file with known answers."""
    sys.path.insert(0, str(ROOT / "tests" / "golden"))
    import make_abi_closure as m

    c = tmp_path / "a.c"
    c.write_text(
        '#include "a.h"\n'
        "static int helper(int x) { return x; }\n"
        "int g_count;\n"
        "static double s_val = 2.0;\n"
        "void (*table[2])(int) = {f0, f1};\n"
        "double f_used(const double *p,\n int n) { /* f_comment() */ return helper(n) + g2(p); }\n"
        "int proto_only(int);\n"
        "#if A\nint alt(void) { return 1; }\n#else\nint alt(void) { return 2; }\n#endif\n"
    )
    d = m.definitions(c, variables=True)
    assert d == {"helper": True, "g_count": False, "s_val": True, "table": False, "f_used": False, "alt": False}
    w = m.body_words(m.strip_c(c.read_text()))
    assert w.get("g2") == 1 and w.get("f0") == 1 and w.get("f1") == 1 and w.get("helper") == 1
    assert "f_comment" not in w and "proto_only" not in w and "f_used" not in w


def test_rw_iterate_visits_sites_in_ildg_order():
    """io.c:553-614 on a host-only lattice with a permuted direction map:
    the callback sequence against an enumeration written from the file
    format (t, z, y, x, x fastest; directions x, y, z, t)."""
    if not LIB.exists():
        pytest.skip("library not built")
    import openstaple_amd as ost
    from openstaple_amd import lib

    nd = (4, 6, 2, 8)  # machine extents d0..d3
    assert lib.ostp_set_lattice_host(*nd, 0) == 0
    # direction map of the host-only lattice: identity (the device context is not initialised here)
    calls = []

    def cb(idxh, parity, dirm, data, swap, fp):
        calls.append((idxh, parity, dirm, swap))

    fn = ost.RW_SITE_FN(cb)
    lib.rw_iterate_on_global_sites_lx_xyzt_axis_ordering(fn, None, None, 0)
    exp = []
    nx, ny, nz, nt = nd
    for t in range(nt):
        for z in range(nz):
            for y in range(ny):
                for x in range(nx):
                    idxh = (x + nx * (y + ny * (z + nz * t))) // 2
                    for dr in range(4):
                        exp.append((idxh, (x + y + z + t) % 2, dr, 1))
    assert calls == exp
    calls.clear()
    lib.rw_iterate_on_global_sites_lx_xyzt_axis_ordering(fn, None, None, 1)
    exp = []
    for t in range(nt):
        for z in range(nz):
            for y in range(ny):
                for xt in range(nx // 2):
                    x = 2 * xt + ((y + z + t) & 1)
                    exp.append(((x + nx * (y + ny * (z + nz * t))) // 2, 0, 0, 1))
    assert calls == exp
    # every even half-site exactly once
    assert sorted(c[0] for c in calls) == list(range(nx * ny * nz * nt // 2))
