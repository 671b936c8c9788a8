"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
Buffers are numpy arrays in the reference layout (see openstaple_amd.Lattice).
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
LIB_PATH = ORACLE_DIR / "liboracle.so"


def build_oracle() -> pathlib.Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "liboracle.so"], check=True)
    return LIB_PATH


if not LIB_PATH.exists():
    build_oracle()

lib = C.CDLL(str(LIB_PATH))


class Geom(C.Structure):
    _fields_ = [("nd", C.c_int * 4), ("sizeh", C.c_long), ("par_temp", C.c_int)]


class Tricks(C.Structure):
    _fields_ = [("restarting_every", C.c_int), ("mixed_precision_delta", C.c_double)]


class RatApprox(C.Structure):
    _fields_ = [
        ("exponent_num", C.c_int), ("exponent_den", C.c_int), ("approx_order", C.c_int),
        ("lambda_min", C.c_double), ("lambda_max", C.c_double), ("gmp_remez_precision", C.c_int),
        ("error", C.c_double), ("RA_a0", C.c_double),
        ("RA_a", C.c_double * 25), ("RA_b", C.c_double * 25),
    ]


_P, _I, _D = C.c_void_p, C.c_int, C.c_double
_G = C.POINTER(Geom)
for name, res, args in [
    ("or_geom_init", None, [_G, _I, _I, _I, _I, _I]),
    ("or_snum", _I, [_G, _I, _I, _I, _I]),
    ("or_nn_tables", None, [_G, _P, _P]),
    ("or_calc_u1_phases", None, [_G, _P, _P, _D, _D, _P]),
    ("or_phases_to_float", None, [_G, _P, _P]),
    ("or_acc_Deo", None, [_G, _P, _P, _P, _P]),
    ("or_acc_Doe", None, [_G, _P, _P, _P, _P]),
    ("or_acc_Deo_f", None, [_G, _P, _P, _P, _P]),
    ("or_acc_Doe_f", None, [_G, _P, _P, _P, _P]),
    ("or_fermion_matrix_multiplication_shifted", None, [_G, _P, _P, _P, _P, _P, _D, _D]),
    ("or_l2norm2", _D, [_G, _P]),
    ("or_real_scal_prod", _D, [_G, _P, _P]),
    ("or_ker_invert", _I, [_G, _P, _P, _D, _P, _P, _D, _P, _P, _P, _P, _I, _D, C.POINTER(Tricks), C.POINTER(_I)]),
    ("or_inverter_mixed_precision", _I,
     [_G, _P, _P, _P, _P, _D, _P, _P, _D, _I, _D, C.POINTER(Tricks), C.POINTER(_I)]),
    ("or_multishift_invert", _I, [_G, _P, _P, _D, _I, _P, _P, _P, _D, _P, _P, _P, _P, _P, _I, C.POINTER(_I)]),
    ("or_recombine", None, [_G, _P, _P, _P, _I, _D, _P]),
    ("or_vec3_d2f", None, [_G, _P, _P]),
    ("or_su3_d2f", None, [_G, _P, _P, _I]),
    ("or_philox4x32_10", None, [_P, _P, _P]),
    ("or_gen_vec3_gauss", None, [_G, _P, C.c_ulonglong, C.c_uint]),
    ("or_gen_vec3_z2", None, [_G, _P, C.c_ulonglong, C.c_uint]),
    ("or_gen_conf_cold", None, [_G, _P, _D, C.c_ulonglong]),
    ("or_gen_conf_expH", None, [_G, _P, _D, C.c_ulonglong]),
    ("or_find_max_eigenvalue", _D, [_G, _P, _P, _D, _P, _P, _P, C.POINTER(_I)]),
    ("or_find_min_eigenvalue", _D, [_G, _P, _P, _D, _P, _P, _P, _D, C.POINTER(_I)]),
    ("or_ratapprox_read", _I, [C.POINTER(RatApprox), C.c_char_p]),
    ("or_ratapprox_evaluate", _D, [C.POINTER(RatApprox), _D]),
    ("or_ratapprox_rescale", None, [C.POINTER(RatApprox), C.POINTER(RatApprox), _P]),
]:
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


def p(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


class Oracle:
    """Oracle bound to one lattice shape."""

    def __init__(self, nd0, nd1, nd2, nd3, par_temp=False):
        self.g = Geom()
        lib.or_geom_init(C.byref(self.g), nd0, nd1, nd2, nd3, int(par_temp))
        self.sizeh = self.g.sizeh
        self.nd = (nd0, nd1, nd2, nd3)
        self.par_temp = par_temp

    @property
    def gp(self):
        return C.byref(self.g)

    # containers in reference layout
    def vec3(self, single=False):
        return np.zeros((3, self.sizeh), dtype=np.complex64 if single else np.complex128)

    def gauge(self, single=False):
        per = (72 + (4 if self.par_temp else 0)) if single else (144 + (8 if self.par_temp else 0))
        return np.zeros(8 * per * self.sizeh, dtype=np.uint8)

    def phases(self, map4=(0, 1, 2, 3), bf=(0, 0, 0, 0, 0, 0), mu=0.0, charge=0.0):
        ph = np.zeros((8, self.sizeh))
        m = (C.c_int * 4)(*map4)
        b = (C.c_double * 6)(*bf)
        lib.or_calc_u1_phases(self.gp, m, b, mu, charge, p(ph))
        return ph

    def nn_tables(self):
        nnp = np.zeros((self.sizeh, 4, 2), dtype=np.int32)
        nnm = np.zeros((self.sizeh, 4, 2), dtype=np.int32)
        lib.or_nn_tables(self.gp, p(nnp), p(nnm))
        return nnp, nnm

    def deo(self, u, inp, ph):
        out = self.vec3()
        lib.or_acc_Deo(self.gp, p(u), p(out), p(inp), p(ph))
        return out

    def doe(self, u, inp, ph):
        out = self.vec3()
        lib.or_acc_Doe(self.gp, p(u), p(out), p(inp), p(ph))
        return out

    def deo_f(self, u_f, inp, ph_f):
        out = self.vec3(single=True)
        lib.or_acc_Deo_f(self.gp, p(u_f), p(out), p(inp), p(ph_f))
        return out

    def doe_f(self, u_f, inp, ph_f):
        out = self.vec3(single=True)
        lib.or_acc_Doe_f(self.gp, p(u_f), p(out), p(inp), p(ph_f))
        return out

    def mdagm(self, u, inp, ph, mass, shift=0.0):
        out, tmp = self.vec3(), self.vec3()
        lib.or_fermion_matrix_multiplication_shifted(self.gp, p(u), p(out), p(inp), p(tmp), p(ph), mass, shift)
        return out

    def cg(self, u, ph, mass, b, res, max_cg=10000, shift=0.0, restart=10000, x0=None):
        x = self.vec3() if x0 is None else x0.copy()
        r, h, s, pp = (self.vec3() for _ in range(4))
        tr = Tricks(restart, 0.1)
        cg = C.c_int(0)
        st = lib.or_ker_invert(self.gp, p(u), p(ph), mass, p(x), p(b), res, p(r), p(h), p(s), p(pp), max_cg, shift,
                               C.byref(tr), C.byref(cg))
        return st, cg.value, x

    def find_max_eigenvalue(self, u, ph, mass, start):
        """Power method of find_min_max.c:21-58 from `start`; (max, iterations)."""
        r, h, pp = self.vec3(), self.vec3(), start.copy()
        it = C.c_int(0)
        v = lib.or_find_max_eigenvalue(self.gp, p(u), p(ph), mass, p(r), p(h), p(pp), C.byref(it))
        return v, it.value

    def find_min_eigenvalue(self, u, ph, mass, start, vmax):
        r, h, pp = self.vec3(), self.vec3(), start.copy()
        it = C.c_int(0)
        v = lib.or_find_min_eigenvalue(self.gp, p(u), p(ph), mass, p(r), p(h), p(pp), vmax, C.byref(it))
        return v, it.value

    def eo_inversion(self, u, ph, mass, b_e, b_o, res, max_cg=10000):
        x_e, x_o, phi_e, phi_o, r, h, s, pp = (self.vec3() for _ in range(8))
        tr = Tricks(10000, 0.1)
        cg = C.c_int(0)
        st = lib.or_eo_inversion(self.gp, p(u), p(ph), mass, res, max_cg, p(b_e), p(b_o), p(x_e), p(x_o), p(phi_e),
                                 p(phi_o), p(r), p(h), p(s), p(pp), C.byref(tr), C.byref(cg))
        return st, cg.value, x_e, x_o

    # ---- parallel-transport preconditioner (oracle/ptc.c); w: (C, 8) complex
    def ptc_transports(self, u, ph, r):
        t = np.zeros((8, 3, self.sizeh), dtype=np.complex128)
        lib.or_ptc_transports(self.gp, p(u), p(ph), p(np.ascontiguousarray(r)), p(t))
        return t

    def ptc_apply(self, u, ph, w0, w, r):
        w = np.ascontiguousarray(w, dtype=np.complex128)
        z = self.vec3()
        lib.or_ptc_apply(self.gp, p(u), p(ph), w.shape[0], w0, p(w), p(np.ascontiguousarray(r)), p(z))
        return z

    def ptc_loss_grad(self, u, ph, w0, w, b, x):
        w = np.ascontiguousarray(w, dtype=np.complex128)
        g = np.zeros(1 + 2 * w.size)
        L = lib.or_ptc_loss_grad(self.gp, p(u), p(ph), w.shape[0], w0, p(w), p(np.ascontiguousarray(b)),
                                 p(np.ascontiguousarray(x)), p(g))
        return L, g

    def init_k(self, conf, c_r, def_axis, def_vec, map4=(0, 1, 2, 3), defect_info_config=0):
        """K planes of `conf` (par_temp layout, uint8 view) set in place; returns
        the 100-int defect_info (GAUGE_ACT_TLSM layout)."""
        assert self.par_temp
        out = np.zeros(100, dtype=np.int32)
        lib.or_init_k(self.gp, (C.c_int * 4)(*map4), p(conf), c_r, def_axis, (C.c_int * 3)(*def_vec), p(out),
                      defect_info_config)
        return out

    def k_planes(self, conf):
        """(8, sizeh) view of the K planes of a par_temp gauge field."""
        assert self.par_temp
        per = 152 * self.sizeh
        return np.stack([conf[l * per + 144 * self.sizeh:(l + 1) * per].view(np.float64) for l in range(8)])

    def mixed(self, u, u_f, ph, ph_f, mass, b, res, max_cg=10000, shift=0.0, delta=0.1, x0=None):
        x = self.vec3() if x0 is None else x0.copy()
        tr = Tricks(10000, delta)
        cg = C.c_int(0)
        st = lib.or_inverter_mixed_precision(self.gp, p(u), p(u_f), p(ph), p(ph_f), mass, p(x), p(b), res, max_cg,
                                             shift, C.byref(tr), C.byref(cg))
        return st, cg.value, x

    def multishift(self, u, ph, mass, shifts, b, res, max_cg=10000):
        order = len(shifts)
        out = np.zeros((max(order, 1), 3, self.sizeh), dtype=np.complex128)
        sf = np.zeros_like(out)
        r, h, s, pp = (self.vec3() for _ in range(4))
        sh = np.ascontiguousarray(shifts, dtype=np.float64)
        cg = C.c_int(0)
        st = lib.or_multishift_invert(self.gp, p(u), p(ph), mass, order, p(sh), p(out), p(b), res, p(r), p(h),
                                      p(s), p(pp), p(sf), max_cg, C.byref(cg))
        return st, cg.value, out[:order]

    def recombine(self, xs, b, a0, a):
        out = self.vec3()
        aa = np.ascontiguousarray(a, dtype=np.float64)
        xs = np.ascontiguousarray(xs)
        lib.or_recombine(self.gp, p(xs), p(b), p(out), len(a), a0, p(aa))
        return out

    def to_float_gauge(self, u):
        uf = self.gauge(single=True)
        lib.or_su3_d2f(self.gp, p(u), p(uf), 8)
        return uf

    def gen_gauss(self, seed, stream=0):
        v = self.vec3()
        lib.or_gen_vec3_gauss(self.gp, p(v), seed, stream)
        return v

    def gen_z2(self, seed, stream=0):
        v = self.vec3()
        lib.or_gen_vec3_z2(self.gp, p(v), seed, stream)
        return v

    def gen_cold(self, eps, seed):
        u = self.gauge()
        lib.or_gen_conf_cold(self.gp, p(u), eps, seed)
        return u

    def gen_expH(self, beta, seed):
        u = self.gauge()
        lib.or_gen_conf_expH(self.gp, p(u), beta, seed)
        return u

    def links(self, u):
        """(8,3,3,sizeh) complex view (par_temp False only)."""
        assert not self.par_temp
        return u.view(np.complex128).reshape(8, 3, 3, self.sizeh)


PT_ACTION = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_void_p)
PT_UNIFORM = C.CFUNCTYPE(C.c_double, C.c_void_p)
lib.or_manage_replica_swaps.argtypes = [_I, _I, _P, _P, PT_ACTION, PT_UNIFORM, _P, C.POINTER(_I), _P, _P]


def manage_replica_swaps(n, is_evenodd, cr_vec, label, action, uniform, counters):
    """Oracle label manager; label (int32 array) and counters = (swap_num
    c_int, all_swap_vet int32[n], acceptance_vet int32[n]) updated in place."""
    swap_num, all_swap, acc = counters
    cr = np.ascontiguousarray(cr_vec, dtype=np.float64)
    lib.or_manage_replica_swaps(n, int(is_evenodd), p(cr), p(label), action, uniform, None, C.byref(swap_num),
                                p(all_swap), p(acc))


def metro_swap(ds, u):
    return lib.or_metro_swap(C.c_double(ds), C.c_double(u))


def philox(ctr, key):
    c = (C.c_uint * 4)(*ctr)
    k = (C.c_uint * 2)(*key)
    o = (C.c_uint * 4)()
    lib.or_philox4x32_10(c, k, o)
    return list(o)


def read_remez(path) -> RatApprox:
    ra = RatApprox()
    rc = lib.or_ratapprox_read(C.byref(ra), str(path).encode())
    if rc != 0:
        raise ValueError(f"or_ratapprox_read({path}) -> {rc}")
    return ra


# ---- gauge part (oracle/gauge.c)
C_ZERO_TLSM = 5.0 * 0.33333333333333333333333  # common_defines.h:71-76
for name, res, args in [
    ("or_calc_loc_staples", None, [_G, _P, _P, _D]),
    ("or_rho_times_conf_times_staples_ta", None, [_G, _P, _P, _P, _D, _D]),
    ("or_exp_minus_QA_times_conf", None, [_G, _P, _P, _P]),
    ("or_stout_isotropic", None, [_G, _P, _P, _D, _D]),
    ("or_compute_lambda", None, [_G, _P, _P, _P, _P]),
    ("or_compute_sigma", None, [_G, _P, _P, _P, _P, _D]),
    ("or_sigma_prime_to_sigma", None, [_G, _P, _P, _P, _P, _D, _D]),
    ("or_unitarize_conf", None, [_G, _P]),
    ("or_direct_product_into_auxmat", None, [_G, _P, _P, _P, _D]),
    ("or_multiply_backfield_times_force", None, [_G, _P, _P, _P]),
    ("or_conf_times_force_ta", None, [_G, _P, _P, _P]),
    ("or_eo_inversion", _I, [_G, _P, _P, _D, _D, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.POINTER(Tricks),
                             C.POINTER(_I)]),
    ("or_init_k", None, [_G, _P, _P, _D, _I, _P, _P, _I]),
    ("or_ptc_transports", None, [_G, _P, _P, _P, _P]),
    ("or_ptc_adjoint_add", None, [_G, _P, _P, _P, _P]),
    ("or_ptc_apply", None, [_G, _P, _P, _I, _D, _P, _P, _P]),
    ("or_ptc_loss_grad", _D, [_G, _P, _P, _I, _D, _P, _P, _P, _P]),
    ("or_metro_swap", _I, [_D, _D]),
]:
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


def _ta_zero(o):
    return np.zeros(8 * 64 * o.sizeh, dtype=np.uint8)


def ta_view(o, t):
    """(c (8,3,sizeh) complex, d (8,2,sizeh) real) views of a tamat/thmat soa[8]."""
    raw = t.view(np.uint8).reshape(8, 64 * o.sizeh)
    c = raw[:, : 48 * o.sizeh].view(np.complex128).reshape(8, 3, o.sizeh)
    d = raw[:, 48 * o.sizeh:].view(np.float64).reshape(8, 2, o.sizeh)
    return c, d


def _bind_gauge(cls):
    def staples(self, u, c_zero=C_ZERO_TLSM):
        st = self.gauge()
        lib.or_calc_loc_staples(self.gp, p(u), p(st), c_zero)
        return st

    def qa(self, u, st, rho, c_zero=C_ZERO_TLSM):
        q = _ta_zero(self)
        lib.or_rho_times_conf_times_staples_ta(self.gp, p(u), p(st), p(q), rho, c_zero)
        return q

    def stout(self, u, rho, c_zero=C_ZERO_TLSM):
        out = self.gauge()
        lib.or_stout_isotropic(self.gp, p(u), p(out), rho, c_zero)
        return out

    def sigma_prime_to_sigma(self, sigma, u, rho, c_zero=C_ZERO_TLSM):
        s = sigma.copy()
        lam, q = _ta_zero(self), _ta_zero(self)
        lib.or_sigma_prime_to_sigma(self.gp, p(s), p(lam), p(q), p(u), rho, c_zero)
        return s, lam, q

    def unitarize(self, u):
        v = u.copy()
        lib.or_unitarize_conf(self.gp, p(v))
        return v

    def direct_product(self, s, h, aux, a):
        lib.or_direct_product_into_auxmat(self.gp, p(s), p(h), p(aux), a)

    def backfield_times_force(self, ph, aux, out):
        lib.or_multiply_backfield_times_force(self.gp, p(ph), p(aux), p(out))

    def conf_times_force_ta(self, u, aux, ipdot):
        lib.or_conf_times_force_ta(self.gp, p(u), p(aux), p(ipdot))

    for fn in (staples, qa, stout, sigma_prime_to_sigma, unitarize, direct_product, backfield_times_force,
               conf_times_force_ta):
        setattr(cls, fn.__name__, fn)
    cls.ta_zero = _ta_zero
    return cls


_bind_gauge(Oracle)
