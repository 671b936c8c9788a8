"""CPU baseline of the fermion hot path -- the oracle restatement timed on the
host of the GPU machine (test infrastructure, used by bench.py only).

build(dir): compiles oracle/*.c twice with gcc -O3 -march=native on this host
            (serial, and OpenMP over the site loops of the Dslash and the
            BLAS); call it before the process touches the GPU (it runs make).
run(...):   times, for each build, the workloads of the CPU-baseline plan:
            * the Dslash (Doe + Deo pairs) at 8^4 and at the bench lattice,
              effective GB/s under the 928 B / half-site model;
            * the 8^4 plain FP64 CG: exp(i 0.3 H) links (seed 1), antiperiodic
              time, Gaussian even source, m = 0.05, to 1e-10 (seconds and
              iterations);
            * a few CG-M iterations at the bench lattice with the 19 shifts of
              the bench's rational approximation, seconds per iteration.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import platform
import subprocess
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
BYTES_PER_SITE_F64 = 928


def build(outdir: str) -> dict:
    """Build the two libraries into outdir; returns their paths (or an error)."""
    r = subprocess.run(["make", "-s", "-j2", "-C", str(HERE), "bench", f"BENCHDIR={outdir}"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        return {"error": (r.stderr or r.stdout)[-400:]}
    return {"serial": os.path.join(outdir, "liboracle_serial.so"), "omp": os.path.join(outdir, "liboracle_omp.so")}


class _Geom(C.Structure):
    _fields_ = [("nd", C.c_int * 4), ("sizeh", C.c_long), ("par_temp", C.c_int)]


class _Tricks(C.Structure):
    _fields_ = [("restarting_every", C.c_int), ("mixed_precision_delta", C.c_double)]


def _bind(path):
    lib = C.CDLL(path)
    P, I, D, G = C.c_void_p, C.c_int, C.c_double, C.POINTER(_Geom)
    for name, res, args in [
        ("or_geom_init", None, [G, I, I, I, I, I]),
        ("or_calc_u1_phases", None, [G, P, P, D, D, P]),
        ("or_acc_Deo", None, [G, P, P, P, P]),
        ("or_acc_Doe", None, [G, P, P, P, P]),
        ("or_ker_invert", I, [G, P, P, D, P, P, D, P, P, P, P, I, D, C.POINTER(_Tricks), C.POINTER(I)]),
        ("or_multishift_invert", I, [G, P, P, D, I, P, P, P, D, P, P, P, P, P, I, C.POINTER(I)]),
        ("or_gen_vec3_gauss", None, [G, P, C.c_ulonglong, C.c_uint]),
        ("or_gen_conf_cold", None, [G, P, D, C.c_ulonglong]),
        ("or_gen_conf_expH", None, [G, P, D, C.c_ulonglong]),
    ]:
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    return lib


def _ptr(a):
    return a.ctypes.data


class _Lat:
    def __init__(self, lib, shape):
        self.lib, self.g = lib, _Geom()
        lib.or_geom_init(C.byref(self.g), *shape, 0)
        self.sizeh = self.g.sizeh
        self.u = np.zeros(8 * 144 * self.sizeh, dtype=np.uint8)
        self.ph = np.zeros((8, self.sizeh))
        lib.or_calc_u1_phases(C.byref(self.g), (C.c_int * 4)(0, 1, 2, 3), (C.c_double * 6)(), 0.0, 0.0,
                              _ptr(self.ph))

    def vec(self):
        return np.zeros((3, self.sizeh), dtype=np.complex128)

    def gauss(self, seed):
        v = self.vec()
        self.lib.or_gen_vec3_gauss(C.byref(self.g), _ptr(v), seed, 0)
        return v


def _dslash_gbs(lat, seconds, min_pairs=2):
    lib, g = lat.lib, C.byref(lat.g)
    psi, h, out = lat.gauss(2), lat.vec(), lat.vec()
    n = 0
    t0 = time.perf_counter()
    while True:
        lib.or_acc_Doe(g, _ptr(lat.u), _ptr(h), _ptr(psi), _ptr(lat.ph))
        lib.or_acc_Deo(g, _ptr(lat.u), _ptr(out), _ptr(h), _ptr(lat.ph))
        n += 1
        if n >= min_pairs and time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    return {"pairs": n, "seconds": round(dt, 3), "s_per_dslash": dt / (2 * n),
            "GBs": round(2 * n * BYTES_PER_SITE_F64 * lat.sizeh / dt / 1e9, 3)}


def _cg_config1(lib):
    lat = _Lat(lib, (8, 8, 8, 8))
    lib.or_gen_conf_expH(C.byref(lat.g), _ptr(lat.u), 0.3, 1)
    b = lat.gauss(101)
    x = lat.vec()
    scratch = [lat.vec() for _ in range(4)]
    cg = C.c_int(0)
    tr = _Tricks(10000, 0.1)
    t0 = time.perf_counter()
    st = lib.or_ker_invert(C.byref(lat.g), _ptr(lat.u), _ptr(lat.ph), 0.05, _ptr(x), _ptr(b), 1e-10,
                           *(_ptr(s) for s in scratch), 10000, 0.0, C.byref(tr), C.byref(cg))
    return {"seconds": round(time.perf_counter() - t0, 4), "iterations": cg.value, "status": st}


def _cgm_per_iteration(lat, shifts, iters):
    """Seconds per CG-M iteration: runs cut at 1 and at 1 + iters iterations
    (each ends with the reference's per-shift residual check, which the
    difference cancels)."""
    lib = lat.lib
    lib.or_gen_conf_cold(C.byref(lat.g), _ptr(lat.u), 0.5, 1)
    b = lat.gauss(3)
    order = len(shifts)
    out = np.zeros((order, 3, lat.sizeh), dtype=np.complex128)
    sf = np.zeros_like(out)
    scratch = [lat.vec() for _ in range(4)]
    sh = np.ascontiguousarray(shifts, dtype=np.float64)
    cg = C.c_int(0)
    times = []
    for n in (1, 1 + iters):
        t0 = time.perf_counter()
        lib.or_multishift_invert(C.byref(lat.g), _ptr(lat.u), _ptr(lat.ph), 0.01, order, _ptr(sh), _ptr(out),
                                 _ptr(b), 1e-10, *(_ptr(s) for s in scratch), _ptr(sf), n, C.byref(cg))
        times.append(time.perf_counter() - t0)
    return {"iterations_timed": iters, "s_per_iteration": round((times[1] - times[0]) / iters, 4)}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def run(libs: dict, shape, shifts, dslash_seconds=4.0, cgm_iters=3) -> dict:
    """Time both builds; the result dict is bench.py's cpu_baseline entry."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    out = {"cpu_model": _cpu_model(), "nproc": os.cpu_count(), "omp_threads": threads,
           "compiler": "gcc -O3 -march=native (built on this host)"}
    if "error" in libs:
        out["error"] = libs["error"]
        return out
    for kind in ("serial", "omp"):
        lib = _bind(libs[kind])
        big = _Lat(lib, shape)
        lib.or_gen_conf_cold(C.byref(big.g), _ptr(big.u), 0.5, 1)
        small = _Lat(lib, (8, 8, 8, 8))
        lib.or_gen_conf_cold(C.byref(small.g), _ptr(small.u), 0.5, 1)
        out[kind] = {
            "cores": 1 if kind == "serial" else threads,
            "dslash_8x8x8x8": _dslash_gbs(small, 1.0, min_pairs=20),
            "dslash_%s" % "x".join(map(str, shape)): _dslash_gbs(big, dslash_seconds),
            "cg_8x8x8x8_m0.05_1e-10": _cg_config1(lib),
            "cgm_%s_%d_shifts" % ("x".join(map(str, shape)), len(shifts)): _cgm_per_iteration(big, shifts, cgm_iters),
        }
    return out
