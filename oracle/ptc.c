/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Gauge-equivariant parallel-transport preconditioner for the even-odd
 * staggered system A = m^2 - Deo Doe (no reference counterpart: the learned
 * preconditioner is an extension of the hot path).  FP64 throughout.
 *
 * Transports from even to odd sites, for an odd site y (mu = 0..3):
 *   T_mu     r (y) = U_mu(y) e^{i th_mu(y)} r(y + mu)                 (forward)
 *   T_{mu+4} r (y) = U_mu(y - mu)^dag e^{-i th_mu(y - mu)} r(y - mu)   (backward)
 * (the eight terms of D_oe of OpenAcc/fermion_matrix.c:101-157 before the
 * 1/2 and the signs: D_oe = 1/2 (sum_fwd - sum_bwd)).
 *
 * Network (2 layers, C channels, weights w[c][nu] complex, w0 real):
 *   layer 1 (even -> odd):  h_c = B_c r = sum_nu w[c][nu] T_nu r
 *   layer 2 (odd -> even):  z   = w0 r + sum_c B_c^dag h_c
 * so N = w0 + sum_c B_c^dag B_c is Hermitian, and positive for w0 > 0, as a
 * conjugate-gradient preconditioner must be.
 *
 * Loss over one pair: L = |N b - x|^2.  With e = N b - x, t_nu = T_nu b,
 * a_nu = T_nu e and P[nu][mu] = <a_nu, t_mu> (sum over odd sites, colours):
 *   dL/dw0 = 2 Re <e, b>
 *   dL/dRe w[c][nu] + i dL/dIm w[c][nu] = 2 conj(S[c][nu]),
 *   S[c][nu] = sum_mu conj(w[c][mu]) (conj(P[nu][mu]) + P[mu][nu]).
 * Pins: finite differences of L, Hermiticity and positivity of N, gauge
 * covariance, and the closed form w[0][nu] = s_nu (s = +1 forward, -1
 * backward), other channels 0: N = w0 - 4 Deo Doe (tests/test_oracle_ptc.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline or_dc ld(const or_dc *U, long sizeh, int r, int c, long i) { return U[(3 * r + c) * sizeh + i]; }

/* full 3x3 of a link (rows 0,1 stored, row 2 = conj(row0 x row1)) */
static void link3(const void *u, size_t su3, int l, long sizeh, long i, or_dc m[3][3]) {
    const or_dc *U = (const or_dc *)((const char *)u + l * su3);
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 3; c++) m[r][c] = ld(U, sizeh, r, c, i);
    m[2][0] = conj(m[0][1] * m[1][2] - m[0][2] * m[1][1]);
    m[2][1] = conj(m[0][2] * m[1][0] - m[0][0] * m[1][2]);
    m[2][2] = conj(m[0][0] * m[1][1] - m[0][1] * m[1][0]);
}

/* site coordinates of half-site i of parity par */
static void coords(const or_geom *g, long i, int par, int d[4]) {
    const int nd0h = g->nd[0] / 2;
    long t = i / nd0h;
    const int hd0 = (int)(i - t * nd0h);
    long t2 = t / g->nd[1];
    d[1] = (int)(t - t2 * g->nd[1]);
    d[3] = (int)(t2 / g->nd[2]);
    d[2] = (int)(t2 - (long)d[3] * g->nd[2]);
    d[0] = 2 * hd0 + ((d[1] + d[2] + d[3] + par) & 1);
}

static long shifted(const or_geom *g, const int d[4], int mu, int s) {
    int e[4] = {d[0], d[1], d[2], d[3]};
    e[mu] = (e[mu] + s + g->nd[mu]) % g->nd[mu];
    return or_snum(g, e[0], e[1], e[2], e[3]);
}

/* out[nu] (odd, 8 fields of 3 x sizeh) = T_nu in (in: even field) */
void or_ptc_transports(const or_geom *g, const void *u, const double *ph, const void *in, void *out) {
    const long sh = g->sizeh;
    const size_t su3 = or_su3_bytes(g);
    const or_dc *r = (const or_dc *)in;
    or_dc *t = (or_dc *)out;
    for (long y = 0; y < sh; y++) {
        int d[4];
        coords(g, y, 1, d);
        for (int mu = 0; mu < 4; mu++) {
            or_dc m[3][3];
            /* forward: link 2mu+1 (odd) at y, r(y+mu) */
            const long yp = shifted(g, d, mu, +1);
            link3(u, su3, 2 * mu + 1, sh, y, m);
            const double a = ph[(2 * mu + 1) * sh + y];
            const or_dc e = cos(a) + I * sin(a);
            for (int c = 0; c < 3; c++) {
                or_dc s = 0;
                for (int k = 0; k < 3; k++) s += m[c][k] * r[k * sh + yp];
                t[(mu * 3 + c) * sh + y] = s * e;
            }
            /* backward: link 2mu (even) at y-mu, dagger, r(y-mu) */
            const long ym = shifted(g, d, mu, -1);
            link3(u, su3, 2 * mu, sh, ym, m);
            const double b = ph[(2 * mu) * sh + ym];
            const or_dc eb = cos(b) - I * sin(b);
            for (int c = 0; c < 3; c++) {
                or_dc s = 0;
                for (int k = 0; k < 3; k++) s += conj(m[k][c]) * r[k * sh + ym];
                t[((mu + 4) * 3 + c) * sh + y] = s * eb;
            }
        }
    }
}

/* out (even) += sum_nu T_nu^dag in[nu] (in: 8 odd fields) */
void or_ptc_adjoint_add(const or_geom *g, const void *u, const double *ph, const void *in, void *out) {
    const long sh = g->sizeh;
    const size_t su3 = or_su3_bytes(g);
    const or_dc *h = (const or_dc *)in;
    or_dc *z = (or_dc *)out;
    for (long x = 0; x < sh; x++) {
        int d[4];
        coords(g, x, 0, d);
        for (int mu = 0; mu < 4; mu++) {
            or_dc m[3][3];
            /* adjoint of the forward transport: x = y + mu, y = x - mu (odd), link 2mu+1 at y, dagger */
            const long xm = shifted(g, d, mu, -1);
            link3(u, su3, 2 * mu + 1, sh, xm, m);
            const double a = ph[(2 * mu + 1) * sh + xm];
            const or_dc ea = cos(a) - I * sin(a);
            for (int c = 0; c < 3; c++) {
                or_dc s = 0;
                for (int k = 0; k < 3; k++) s += conj(m[k][c]) * h[(mu * 3 + k) * sh + xm];
                z[c * sh + x] += s * ea;
            }
            /* adjoint of the backward transport: x = y - mu, y = x + mu, link 2mu (even) at x */
            const long xp = shifted(g, d, mu, +1);
            link3(u, su3, 2 * mu, sh, x, m);
            const double b = ph[(2 * mu) * sh + x];
            const or_dc eb = cos(b) + I * sin(b);
            for (int c = 0; c < 3; c++) {
                or_dc s = 0;
                for (int k = 0; k < 3; k++) s += m[c][k] * h[((mu + 4) * 3 + k) * sh + xp];
                z[c * sh + x] += s * eb;
            }
        }
    }
}

/* z = N r; w = C x 8 complex (row c), w0 real */
void or_ptc_apply(const or_geom *g, const void *u, const double *ph, int C, double w0, const or_dc *w,
                  const void *r, void *z) {
    const long sh = g->sizeh, n3 = 3 * sh;
    or_dc *t = malloc(sizeof(or_dc) * 8 * n3), *gm = calloc(8 * n3, sizeof(or_dc));
    or_ptc_transports(g, u, ph, r, t);
    /* g_nu = sum_c conj(w[c][nu]) h_c, h_c = sum_mu w[c][mu] t_mu */
    for (long i = 0; i < n3; i++) {
        for (int c = 0; c < C; c++) {
            or_dc h = 0;
            for (int mu = 0; mu < 8; mu++) h += w[c * 8 + mu] * t[mu * n3 + i];
            for (int nu = 0; nu < 8; nu++) gm[nu * n3 + i] += conj(w[c * 8 + nu]) * h;
        }
    }
    or_dc *zz = (or_dc *)z;
    const or_dc *rr = (const or_dc *)r;
    for (long i = 0; i < n3; i++) zz[i] = w0 * rr[i];
    or_ptc_adjoint_add(g, u, ph, gm, z);
    free(t);
    free(gm);
}

/* L = |N b - x|^2 and its gradient: grad[0] = dL/dw0, grad[1 + 2(8c+nu)] = dL/dRe w[c][nu],
 * grad[2 + 2(8c+nu)] = dL/dIm w[c][nu]. */
double or_ptc_loss_grad(const or_geom *g, const void *u, const double *ph, int C, double w0, const or_dc *w,
                        const void *b, const void *x, double *grad) {
    const long sh = g->sizeh, n3 = 3 * sh;
    or_dc *z = malloc(sizeof(or_dc) * n3), *e = malloc(sizeof(or_dc) * n3);
    or_dc *t = malloc(sizeof(or_dc) * 8 * n3), *a = malloc(sizeof(or_dc) * 8 * n3);
    or_ptc_apply(g, u, ph, C, w0, w, b, z);
    const or_dc *bb = (const or_dc *)b, *xx = (const or_dc *)x;
    double L = 0;
    or_dc s0 = 0;
    for (long i = 0; i < n3; i++) {
        e[i] = z[i] - xx[i];
        L += creal(e[i]) * creal(e[i]) + cimag(e[i]) * cimag(e[i]);
        s0 += conj(e[i]) * bb[i];
    }
    or_ptc_transports(g, u, ph, b, t);
    or_ptc_transports(g, u, ph, e, a);
    or_dc P[8][8];
    for (int nu = 0; nu < 8; nu++)
        for (int mu = 0; mu < 8; mu++) {
            or_dc s = 0;
            for (long i = 0; i < n3; i++) s += conj(a[nu * n3 + i]) * t[mu * n3 + i];
            P[nu][mu] = s;
        }
    grad[0] = 2 * creal(s0);
    for (int c = 0; c < C; c++)
        for (int nu = 0; nu < 8; nu++) {
            or_dc S = 0;
            for (int mu = 0; mu < 8; mu++) S += conj(w[c * 8 + mu]) * (conj(P[nu][mu]) + P[mu][nu]);
            const or_dc G = 2 * conj(S);
            grad[1 + 2 * (8 * c + nu)] = creal(G);
            grad[2 + 2 * (8 * c + nu)] = cimag(G);
        }
    free(z); free(e); free(t); free(a);
    return L;
}
