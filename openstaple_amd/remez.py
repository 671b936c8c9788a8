"""Rational minimax approximations of x^(num/den) for the RHMC, written as the
reference's .REMEZ files.

This is the generator behind the approximation files the RHMC reads: the
reference builds it from RationalApprox/rationalapprox_calc.cc (by tolerance)
and rationalapprox_calc_fixed_order.cc (by order) around a GMP implementation
of the Remez algorithm.  Here it is a multiple-precision Python tool (mpmath),
offline like the reference's: it runs once per approximation, on the host.

Method (the classical Remez exchange for rational functions, relative error):
for f(x) = x^p, p = |num/den|, find R = P/Q with deg P = deg Q = n that
minimises max |R(x)/f(x) - 1| on [lambda_min, lambda_max].  The minimax R is
characterised by 2n + 2 points where the relative error reaches its maximum
E with alternating signs; starting from points spread in log x (a blend of
uniform and Chebyshev spacing; another blend if a start runs into a pole
inside the interval), each iteration
  1. solves P(x_i) = f(x_i) (1 + (-1)^i E) Q(x_i) for the coefficients and E
     (linear once E Q(x_i) is taken from the previous sweep, repeated until
     E settles), in a monomial basis at 3.3 bits per decimal digit of
     working precision;
  2. moves the points to the alternating extrema of the error, located on a
     logarithmic grid and refined by golden-section search in log x;
until the extrema are level to 1e-10 relative.  R is then written in the
partial-fraction form the library evaluates, a0 + sum_k a_k / (x + b_k): the
poles -b_k are the roots of Q and a_k = P(-b_k) / Q'(-b_k).  A negative
exponent uses the inverse Q/P of the approximation of x^p, whose relative
error is the same to first order (the reference's convention: the file's
Error is that of the approximation of x^|p|).

    python -m openstaple_amd.remez TOLERANCE NUM DEN LAMBDA_MIN [ORDER]
    python -m openstaple_amd.remez --order ORDER NUM DEN LAMBDA_MIN
"""
from __future__ import annotations

import argparse
import math
import sys

import mpmath as mp

from . import MAX_APPROX_ORDER, RationalApprox


class RemezError(RuntimeError):
    pass


def _log_points(a, b, n, w):
    """n + 1 points from a to b in log x: a blend (weight w) of uniform and
    Chebyshev spacing"""
    la, lb = mp.log(a), mp.log(b)
    return [mp.exp(la + (lb - la) * ((1 - w) * (1 - mp.cos(mp.pi * i / n)) / 2 + w * mp.mpf(i) / n))
            for i in range(n + 1)]


def _polyval(c, x):
    """sum c[k] x^k"""
    s = mp.mpf(0)
    for ck in reversed(c):
        s = s * x + ck
    return s


class _Remez:
    def __init__(self, p, a, b, n, dps):
        self.p, self.n = mp.mpf(p), n
        self.a, self.b = mp.mpf(a), mp.mpf(b)
        self.dps = dps

    def f(self, x):
        return x ** self.p

    def err(self, P, Q, x):
        return _polyval(P, x) / (self.f(x) * _polyval(Q, x)) - 1

    def solve(self, xs, P, Q, E):
        """P (n+1 coefficients), Q (Q[0] = 1) and E with
        F_i = P(x_i) - f(x_i) (1 + (-1)^i E) Q(x_i) = 0 at the 2n + 2 points,
        by Newton's method from the previous solution"""
        n = self.n
        fx = [self.f(x) for x in xs]
        pw = [[x ** k for k in range(n + 1)] for x in xs]
        tol = mp.mpf(10) ** (-(self.dps // 2))
        for _ in range(40):
            J = mp.matrix(2 * n + 2, 2 * n + 2)
            F = mp.matrix(2 * n + 2, 1)
            for i in range(2 * n + 2):
                s = 1 if i % 2 == 0 else -1
                w = fx[i] * (1 + s * E)
                qv = sum(Q[k] * pw[i][k] for k in range(n + 1))
                F[i] = sum(P[k] * pw[i][k] for k in range(n + 1)) - w * qv
                for k in range(n + 1):
                    J[i, k] = pw[i][k]
                for k in range(1, n + 1):
                    J[i, n + k] = -w * pw[i][k]
                J[i, 2 * n + 1] = -s * fx[i] * qv
            d = mp.lu_solve(J, -F)
            P = [P[k] + d[k] for k in range(n + 1)]
            Q = [Q[0]] + [Q[k] + d[n + k] for k in range(1, n + 1)]
            E = E + d[2 * n + 1]
            if abs(d[2 * n + 1]) <= tol * max(abs(E), tol) and mp.norm(d) <= tol * max(mp.norm(mp.matrix(P)), 1):
                return P, Q, E
        raise RemezError("Newton's method for the levelled error did not converge")

    def extrema(self, P, Q, xs, sub=12):
        """the 2n + 2 alternating extrema of the error: sign segments of the
        error on a logarithmic grid, each maximum refined by golden section"""
        la = [mp.log(x) for x in xs]
        ts = []
        for i in range(len(la) - 1):
            ts += [la[i] + (la[i + 1] - la[i]) * j / sub for j in range(sub)]
        ts.append(la[-1])
        qs = [_polyval(Q, mp.exp(t)) for t in ts]
        if min(qs) <= 0 < max(qs) or min(qs) < 0 <= max(qs):
            raise RemezError("a pole inside the interval")
        es = [self.err(P, Q, mp.exp(t)) for t in ts]
        segs = []  # [sign, index of the largest |e|]
        for j, e in enumerate(es):
            sg = 1 if e >= 0 else -1
            if segs and segs[-1][0] == sg:
                if abs(e) > abs(es[segs[-1][1]]):
                    segs[-1][1] = j
            else:
                segs.append([sg, j])
        want = 2 * self.n + 2
        while len(segs) > want:
            segs.pop(0 if abs(es[segs[0][1]]) < abs(es[segs[-1][1]]) else -1)
        if len(segs) < want:
            if sub < 200:
                return self.extrema(P, Q, xs, 4 * sub)
            raise RemezError("the error has %d alternations, %d expected" % (len(segs), want))
        phi = (mp.sqrt(5) - 1) / 2
        out = []
        for sg, j in segs:
            def g(t, sg=sg):
                return sg * self.err(P, Q, mp.exp(t))

            l, r = ts[max(j - 1, 0)], ts[min(j + 1, len(ts) - 1)]
            c, d = r - phi * (r - l), l + phi * (r - l)
            gc, gd = g(c), g(d)
            for _ in range(70):
                if gc > gd:
                    r, d, gd = d, c, gc
                    c = r - phi * (r - l)
                    gc = g(c)
                else:
                    l, c, gc = c, d, gd
                    d = l + phi * (r - l)
                    gd = g(d)
            cand = [(g(ts[j]), ts[j]), (g((l + r) / 2), (l + r) / 2)]
            if j == 0:
                cand.append((g(ts[0]), ts[0]))
            if j == len(ts) - 1:
                cand.append((g(ts[-1]), ts[-1]))
            v, t = max(cand, key=lambda vt: vt[0])
            out.append((mp.exp(t), sg * v))
        return out

    def run(self, max_iter=80):
        # starting points: half uniform, half Chebyshev in log x; the other
        # spacings if that start leads to a pole inside the interval
        last = None
        for w in (0.25, 0.75, 1.0, 0.0, 0.5):
            try:
                return self._run(_log_points(self.a, self.b, 2 * self.n + 1, w), max_iter)
            except (RemezError, ZeroDivisionError) as e:
                last = e
        raise RemezError("no start converged: %s" % last)

    def _run(self, xs, max_iter):
        n = self.n
        P = [mp.mpf(0)] * (n + 1)
        Q = [mp.mpf(1)] + [mp.mpf(0)] * n
        E = mp.mpf(0)
        for _ in range(max_iter):
            P, Q, E = self.solve(xs, P, Q, E)
            ext = self.extrema(P, Q, xs)
            emax = max(abs(e) for _, e in ext)
            emin = min(abs(e) for _, e in ext)
            xs = [x for x, _ in ext]
            if emax - emin <= mp.mpf("1e-10") * emax:
                return P, Q, abs(E), emax
        raise RemezError("the Remez exchange did not converge in %d iterations" % max_iter)


def _partial_fractions(num, den):
    """a0 + sum_k a_k / (x + b_k) for num(x)/den(x), deg num = deg den"""
    n = len(den) - 1
    roots = mp.polyroots(list(reversed(den)), maxsteps=400, extraprec=4 * mp.mp.prec)
    dden = [k * den[k] for k in range(1, n + 1)]
    a0 = num[n] / den[n]
    res = []
    for r in roots:
        if abs(mp.im(r)) > abs(r) * mp.mpf(10) ** (-(mp.mp.dps // 2)):
            raise RemezError("complex pole %s" % r)
        r = mp.re(r)
        if r >= 0:
            raise RemezError("pole %s on the approximation's side of the axis" % r)
        res.append((-r, _polyval(num, r) / _polyval(dden, r)))
    res.sort(key=lambda t: t[0])
    return a0, [a for _, a in res], [b for b, _ in res]


def generate(exponent_num, exponent_den, lambda_min, order, lambda_max=1.0, dps=None):
    """RationalApprox of x^(num/den) of the given order on [lambda_min, lambda_max]"""
    if order < 1 or order > MAX_APPROX_ORDER:
        raise RemezError("order %d outside 1..%d" % (order, MAX_APPROX_ORDER))
    if exponent_num * exponent_den == 0:
        raise RemezError("exponent %d/%d" % (exponent_num, exponent_den))
    span = math.log10(lambda_max / lambda_min)
    digits = dps or int(40 + order * span)
    with mp.workdps(digits):
        p = mp.mpf(abs(exponent_num)) / abs(exponent_den)
        R = _Remez(p, mp.mpf(lambda_min), mp.mpf(lambda_max), order, digits)
        P, Q, E, _ = R.run()
        if exponent_num * exponent_den > 0:
            a0, a, b = _partial_fractions(P, Q)
        else:
            a0, a, b = _partial_fractions(Q, P)
        ra = RationalApprox()
        ra.exponent_num, ra.exponent_den = exponent_num, exponent_den
        ra.approx_order = order
        ra.lambda_min, ra.lambda_max = float(lambda_min), float(lambda_max)
        ra.gmp_remez_precision = int(digits * 3.33)
        ra.error = float(E)
        ra.RA_a0 = float(a0)
        for k in range(order):
            ra.RA_a[k] = float(a[k])
            ra.RA_b[k] = float(b[k])
        return ra


def generate_to_tolerance(tolerance, exponent_num, exponent_den, lambda_min, order=None, lambda_max=1.0):
    """rationalapprox_calc.cc: from a guessed (or given) order upwards until the
    error is at most tolerance; the guess is the reference's
    ln(tolerance) ln(lambda_min) / 19 + 2"""
    if order is None:
        order = int(math.log(tolerance) * math.log(lambda_min) / 19 + 2)
    while True:
        ra = generate(exponent_num, exponent_den, lambda_min, order, lambda_max)
        if ra.error <= tolerance:
            return ra
        order += 1
        if order > MAX_APPROX_ORDER:
            raise RemezError("tolerance %g needs an order above %d" % (tolerance, MAX_APPROX_ORDER))


def write(path, ra):
    """the .REMEZ text layout of rationalapprox_save (rationalapprox.c:119-143)"""
    with open(path, "w") as fh:
        fh.write("\nApproximation to f(x) = (x)^(%i/%i)\n" % (ra.exponent_num, ra.exponent_den))
        fh.write("Order: %i\n" % ra.approx_order)
        fh.write("Lambda Min: %18.16e\n" % ra.lambda_min)
        fh.write("Lambda Max: %18.16e\n" % ra.lambda_max)
        fh.write("GMP Remez Precision: %i\n" % ra.gmp_remez_precision)
        fh.write("Error: %18.16e\n" % ra.error)
        fh.write("RA_a0 = %18.16e\n" % ra.RA_a0)
        for k in range(ra.approx_order):
            fh.write("RA_a[%d] = %18.16e, RA_b[%d] = %18.16e\n" % (k, ra.RA_a[k], k, ra.RA_b[k]))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--order", type=int, help="fixed order (rationalapprox_calc_fixed_order.cc)")
    ap.add_argument("args", nargs="+", help="[TOLERANCE] NUM DEN LAMBDA_MIN [ORDER]")
    ns = ap.parse_args(argv)
    from . import lib
    if ns.order is not None:
        num, den, lmin = int(ns.args[0]), int(ns.args[1]), float(ns.args[2])
        ra = generate(num, den, lmin, ns.order)
        name = lib.rational_approx_filename_old(ns.order, num, den, lmin)
    else:
        tol, num, den, lmin = float(ns.args[0]), int(ns.args[1]), int(ns.args[2]), float(ns.args[3])
        order = int(ns.args[4]) if len(ns.args) > 4 else None
        ra = generate_to_tolerance(tol, num, den, lmin, order)
        name = lib.rational_approx_filename(tol, num, den, lmin)
    path = C_string(name)
    write(path, ra)
    print("f(x) = x^(%d/%d) on [%g, %g], order %d, error %.6e -> %s" % (
        ra.exponent_num, ra.exponent_den, ra.lambda_min, ra.lambda_max, ra.approx_order, ra.error, path))
    return 0


def C_string(p):
    import ctypes

    s = ctypes.cast(p, ctypes.c_char_p).value.decode()
    ctypes.CDLL(None).free(ctypes.c_void_p(p))
    return s


if __name__ == "__main__":
    sys.exit(main())
