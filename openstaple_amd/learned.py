"""Training and use of the learned parallel-transport preconditioner.

The network and its forward/backward kernels live in the HIP library
(`ostp_ptc_apply`, `ostp_ptc_loss_grad`, `ostp_fcg_solve`,
`ostp_fcg_mixed_solve`); this module holds
the host side: generating (b, x) pairs with the library's plain FP64 CG,
the Adam optimiser over the OSTP_PTC_NPARAM parameters (w0, then Re/Im of
w[c][nu]), and the flexible-CG solve with plain-CG fallback.

Loss: by default the A-norm of the error, <e, A e> with e = N b - x, divided
by <x, A x> = <x, b> (the error norm conjugate gradients minimise); the
Euclidean |e|^2 / |x|^2 is available as `a_norm=False`.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import PTC_CHANNELS, PTC_NPARAM, check, lib


def init_params(lmax: float = 4.0, seed: int = 0, scale: float = 0.01) -> np.ndarray:
    """w0 = 1/lmax (lmax ~ largest eigenvalue of m^2 - Deo Doe), small random w."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PTC_NPARAM)
    p[0] = 1.0 / lmax
    p[1:] = scale * rng.standard_normal(PTC_NPARAM - 1)
    return p


@dataclasses.dataclass
class Operator:
    """Device pointers of one gauge configuration and mass: FP64 links `u`,
    FP32 links `u_f`, and a ferm_param carrying mass, FP64 and FP32 phases."""

    u: int
    u_f: int
    fp: object  # openstaple_amd.FermParam


class Pairs:
    """(b, x) training pairs on the device: b Gaussian (library Philox
    generator, `seed`, streams 0..n-1), x = A^-1 b from the plain FP64 CG."""

    def __init__(self, op: Operator, vec3_bytes: int, n: int, seed: int, res: float = 1e-10, max_cg: int = 40000):
        self.b, self.x, self.iterations = [], [], []
        self.vb = vec3_bytes
        work = [lib.ostp_malloc(vec3_bytes) for _ in range(4)]
        try:
            for k in range(n):
                b, x = lib.ostp_malloc(vec3_bytes), lib.ostp_malloc(vec3_bytes)
                if not b or not x:
                    raise MemoryError("ostp_malloc")
                lib.ostp_generate_vec3_soa_gauss(b, seed, k)
                lib.ostp_memset(x, 0, vec3_bytes)
                cg, tr = C.c_int(0), C.c_double(0)
                st = lib.ostp_cg_solve(op.u, C.byref(op.fp), x, b, res, *work, max_cg, 0.0, 10000, C.byref(cg),
                                       C.byref(tr))
                check(0, "training pair")
                if st != 1:
                    raise RuntimeError(f"plain CG did not converge for pair {k} (true residual {tr.value:.3e})")
                self.b.append(b)
                self.x.append(x)
                self.iterations.append(cg.value)
        finally:
            for w in work:
                lib.ostp_free(w)

    def __len__(self):
        return len(self.b)

    def free(self):
        for p in self.b + self.x:
            lib.ostp_free(p)
        self.b, self.x = [], []


def loss_grad(op: Operator, params: np.ndarray, b: int, x: int, a_norm: bool = True):
    """Loss and gradient of one pair (library kernels)."""
    p32 = np.ascontiguousarray(params, dtype=np.float32)
    g = np.zeros(PTC_NPARAM)
    L = lib.ostp_ptc_loss_grad(p32.ctypes.data, op.u, op.u_f, C.byref(op.fp), b, x, int(a_norm), g.ctypes.data)
    check(0, "ostp_ptc_loss_grad")
    if L < 0:
        raise RuntimeError("ostp_ptc_loss_grad failed")
    return L, g


def norms(op: Operator, pairs: Pairs, a_norm: bool = True):
    """Per-pair normalisation: <x, b> = <x, A x> (A-norm) or |x|^2."""
    out = []
    for b, x in zip(pairs.b, pairs.x):
        out.append(lib.real_scal_prod_global(x, b) if a_norm else lib.l2norm2_global(x))
    return np.array(out)


def train(op: Operator, pairs: Pairs, params: np.ndarray | None = None, steps: int = 300, lr: float = 0.02,
          batch: int = 8, a_norm: bool = True, seed: int = 0, log_every: int = 0):
    """Adam over mini-batches of pairs; returns (params, loss history)."""
    rng = np.random.default_rng(seed)
    theta = init_params() if params is None else np.array(params, dtype=np.float64)
    nrm = norms(op, pairs, a_norm)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    hist = []
    for it in range(1, steps + 1):
        idx = rng.choice(len(pairs), size=min(batch, len(pairs)), replace=False)
        Ls, G = 0.0, np.zeros_like(theta)
        for k in idx:
            L, g = loss_grad(op, theta, pairs.b[k], pairs.x[k], a_norm)
            Ls += L / nrm[k]
            G += g / nrm[k]
        Ls /= len(idx)
        G /= len(idx)
        m = 0.9 * m + 0.1 * G
        v = 0.999 * v + 0.001 * G * G
        theta -= lr * (m / (1 - 0.9 ** it)) / (np.sqrt(v / (1 - 0.999 ** it)) + 1e-12)
        hist.append(Ls)
        if log_every and it % log_every == 0:
            print(f"step {it}: loss {Ls:.5f} w0 {theta[0]:.4f}", flush=True)
    return theta, np.array(hist)


def solve(op: Operator, params: np.ndarray, x: int, b: int, res: float = 1e-10, max_cg: int = 40000):
    """Flexible CG with the network; returns (status, fcg iterations,
    plain-CG fallback iterations, FP64 true residual)."""
    p32 = np.ascontiguousarray(params, dtype=np.float32)
    cg, fb, tr = C.c_int(0), C.c_int(0), C.c_double(0)
    st = lib.ostp_fcg_solve(op.u, C.byref(op.fp), op.u_f, p32.ctypes.data, x, b, res, max_cg, C.byref(cg),
                            C.byref(fb), C.byref(tr))
    check(0, "ostp_fcg_solve")
    return st, cg.value, fb.value, tr.value


def solve_mixed(op: Operator, params: np.ndarray, x: int, b: int, res: float = 1e-10, max_cg: int = 40000,
                inner_res: float = 0.1):
    """Flexible CG with the network in FP32 and FP64 reliable updates at
    every `inner_res` drop of the residual; returns (status, FP32 iterations,
    reliable updates, plain-CG fallback iterations, FP64 true residual)."""
    p32 = np.ascontiguousarray(params, dtype=np.float32)
    cg, outer, fb, tr = C.c_int(0), C.c_int(0), C.c_int(0), C.c_double(0)
    st = lib.ostp_fcg_mixed_solve(op.u, C.byref(op.fp), op.u_f, p32.ctypes.data, x, b, res, max_cg, inner_res,
                                  C.byref(cg), C.byref(outer), C.byref(fb), C.byref(tr))
    check(0, "ostp_fcg_mixed_solve")
    return st, cg.value, outer.value, fb.value, tr.value


__all__ = ["init_params", "Operator", "Pairs", "loss_grad", "norms", "train", "solve", "solve_mixed",
           "PTC_CHANNELS"]
