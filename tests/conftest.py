import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = ROOT / "tests"
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C ABI)")
    config.addinivalue_line("markers", "slow: long CPU test")


REF_DIR = ROOT / "oracle" / "_ref"


def ref_lib(shape):
    """ctypes handle of the reference's own FP64 sources built for `shape`, or None."""
    import ctypes

    p = REF_DIR / ("libref_%dx%dx%dx%d.so" % tuple(shape))
    if not p.exists():
        return None
    return ctypes.CDLL(str(p))
