"""Helpers for the GPU tests: device buffers through the library's own C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np

import openstaple_amd as ost
from openstaple_amd import lib


class Dev:
    """A device copy of a host array (ostp_malloc + memcpy)."""

    def __init__(self, a: np.ndarray | None = None, nbytes: int | None = None):
        self.nbytes = a.nbytes if a is not None else int(nbytes)
        self.p = lib.ostp_malloc(self.nbytes)
        if not self.p:
            raise MemoryError(lib.ostp_last_error())
        if a is not None:
            self.put(a)
        else:
            lib.ostp_memset(self.p, 0, self.nbytes)

    def put(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        ost.check(lib.ostp_memcpy_htod(self.p, a.ctypes.data, self.nbytes), "htod")

    def get(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert out.nbytes == self.nbytes
        ost.check(lib.ostp_memcpy_dtoh(out.ctypes.data, self.p, self.nbytes), "dtoh")
        return out

    def free(self) -> None:
        if self.p:
            lib.ostp_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def init_lattice(shape, par_temp=False) -> ost.Lattice:
    L = ost.Lattice(*shape, par_temp=par_temp)
    L.init()
    return L


def rel(a, b) -> float:
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def cint():
    return C.c_int(0)
