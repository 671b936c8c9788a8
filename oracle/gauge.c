/*
 * CPU oracle, gauge part -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Stout smearing and its force chain, unitarisation and the fermion-force
 * building blocks, restated from
 *   OpenAcc/plaquettes.c:196-259        (staples, x C_ZERO)
 *   OpenAcc/su3_utilities.h:660-977,1097-1152, su3_utilities.c:17-40,210-239
 *   OpenAcc/stouting.c:74-165,171-546,550-1302  (stout, Lambda, Sigma)
 *   OpenAcc/fermion_force.c:52-162
 *   OpenAcc/fermion_force_utilities.{c:30-153, h:16-190}
 * with 3x3 matrix products in place of the reference's unrolled expressions
 * (same mathematics, rounding differs in the last bits).
 *
 * Field layouts: su3_soa[8] as in oracle.h; tamat_soa / thmat_soa are
 * c01,c02,c12 complex planes followed by two double planes (struct_c_def.h:45-58).
 */
#include "oracle.h"
#include "su3.h"

#include <stdlib.h>
#include <string.h>

#define OR_ONE_BY_THREE 0.33333333333333333333333

static inline or_dc *upl(const or_geom *g, const void *u, int link, int r, int c)
{
    return (or_dc *)((char *)u + (size_t)link * or_su3_bytes(g) + (size_t)(3 * r + c) * 16 * g->sizeh);
}
static inline size_t ta_bytes(const or_geom *g) { return (size_t)64 * g->sizeh; }
static inline or_dc *ta_c(const or_geom *g, const void *t, int link, int k)
{
    return (or_dc *)((char *)t + (size_t)link * ta_bytes(g) + (size_t)k * 16 * g->sizeh);
}
static inline double *ta_d(const or_geom *g, const void *t, int link, int k)
{
    return (double *)((char *)t + (size_t)link * ta_bytes(g) + (size_t)48 * g->sizeh + (size_t)k * 8 * g->sizeh);
}

/* rows 0,1 loaded, row 2 rebuilt */
static void load_su3(const or_geom *g, const void *u, int link, long i, or_su3 *m)
{
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 3; c++) m->comp[r][c] = upl(g, u, link, r, c)[i];
    or_rebuild3row(m);
}
/* all 9 entries loaded (gl(3) fields: staples, Sigma) */
static void load_gl3(const or_geom *g, const void *u, int link, long i, or_su3 *m)
{
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) m->comp[r][c] = upl(g, u, link, r, c)[i];
}
static void store_gl3(const or_geom *g, void *u, int link, long i, const or_su3 *m, int rows)
{
    for (int r = 0; r < rows; r++)
        for (int c = 0; c < 3; c++) upl(g, u, link, r, c)[i] = m->comp[r][c];
}
static void load_ta(const or_geom *g, const void *t, int link, long i, or_tamat *q)
{
    q->c01 = ta_c(g, t, link, 0)[i];
    q->c02 = ta_c(g, t, link, 1)[i];
    q->c12 = ta_c(g, t, link, 2)[i];
    q->ic00 = ta_d(g, t, link, 0)[i];
    q->ic11 = ta_d(g, t, link, 1)[i];
}
static void store_ta(const or_geom *g, void *t, int link, long i, const or_tamat *q)
{
    ta_c(g, t, link, 0)[i] = q->c01;
    ta_c(g, t, link, 1)[i] = q->c02;
    ta_c(g, t, link, 2)[i] = q->c12;
    ta_d(g, t, link, 0)[i] = q->ic00;
    ta_d(g, t, link, 1)[i] = q->ic11;
}

static void mul(const or_su3 *a, const or_su3 *b, or_su3 *o)
{
    or_su3 t;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            t.comp[r][c] = a->comp[r][0] * b->comp[0][c] + a->comp[r][1] * b->comp[1][c] + a->comp[r][2] * b->comp[2][c];
    *o = t;
}
static void dag(const or_su3 *a, or_su3 *o)
{
    or_su3 t;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) t.comp[r][c] = conj(a->comp[c][r]);
    *o = t;
}
/* thmat -> full Hermitian 3x3 */
static void th_full(const or_thmat *l, or_su3 *o)
{
    o->comp[0][0] = l->rc00; o->comp[0][1] = l->c01; o->comp[0][2] = l->c02;
    o->comp[1][0] = conj(l->c01); o->comp[1][1] = l->rc11; o->comp[1][2] = l->c12;
    o->comp[2][0] = conj(l->c02); o->comp[2][1] = conj(l->c12); o->comp[2][2] = -l->rc00 - l->rc11;
}
static void load_th(const or_geom *g, const void *t, int link, long i, or_thmat *l)
{
    l->c01 = ta_c(g, t, link, 0)[i];
    l->c02 = ta_c(g, t, link, 1)[i];
    l->c12 = ta_c(g, t, link, 2)[i];
    l->rc00 = ta_d(g, t, link, 0)[i];
    l->rc11 = ta_d(g, t, link, 1)[i];
}
static void store_th(const or_geom *g, void *t, int link, long i, const or_thmat *l)
{
    ta_c(g, t, link, 0)[i] = l->c01;
    ta_c(g, t, link, 1)[i] = l->c02;
    ta_c(g, t, link, 2)[i] = l->c12;
    ta_d(g, t, link, 0)[i] = l->rc00;
    ta_d(g, t, link, 1)[i] = l->rc11;
}

/* per-site iteration with precomputed neighbour tables */
typedef struct {
    int *nnp, *nnm;
} nn_t;
static nn_t nn_make(const or_geom *g)
{
    nn_t n;
    n.nnp = malloc(sizeof(int) * 8 * g->sizeh);
    n.nnm = malloc(sizeof(int) * 8 * g->sizeh);
    or_nn_tables(g, n.nnp, n.nnm);
    return n;
}
static void nn_free(nn_t *n) { free(n->nnp); free(n->nnm); }
#define NNP(n, i, mu, par) ((n).nnp[(size_t)(i) * 8 + (mu) * 2 + (par)])
#define NNM(n, i, mu, par) ((n).nnm[(size_t)(i) * 8 + (mu) * 2 + (par)])

/* ----------------------------------------------------------------- staples */
/* calc_loc_staples_nnptrick_all_onlyferms (plaquettes.c:196-259): for the link
 * (r, mu) the sum over nu != mu of
 *   right: U_nu(r+mu) U_mu(r+nu)^+ U_nu(r)^+
 *   left : U_nu(r+mu-nu)^+ U_mu(r-nu)^+ U_nu(r-nu)
 * each times c_zero, accumulated into stap (full 3x3, row 2 of each product
 * rebuilt from its rows 0,1 as the reference does). */
void or_calc_loc_staples(const or_geom *g, const void *u, void *stap, double c_zero)
{
    nn_t n = nn_make(g);
    memset(stap, 0, 8 * or_su3_bytes(g));
    for (long idxh = 0; idxh < g->sizeh; idxh++)
        for (int parity = 0; parity < 2; parity++)
            for (int mu = 0; mu < 4; mu++) {
                const int dir_link = 2 * mu + parity;
                or_su3 acc;
                memset(&acc, 0, sizeof acc);
                for (int iter = 0; iter < 3; iter++) {
                    static const int perp[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
                    const int nu = perp[mu][iter];
                    const int idx_pmu = NNP(n, idxh, mu, parity);
                    const int idx_pnu = NNP(n, idxh, nu, parity);
                    or_su3 A, B, C, T;
                    load_su3(g, u, 2 * nu + !parity, idx_pmu, &A);
                    load_su3(g, u, 2 * mu + !parity, idx_pnu, &B);
                    load_su3(g, u, 2 * nu + parity, idxh, &C);
                    dag(&B, &B);
                    dag(&C, &C);
                    mul(&A, &B, &T);
                    mul(&T, &C, &T);
                    or_rebuild3row(&T);
                    for (int r = 0; r < 3; r++)
                        for (int c = 0; c < 3; c++) acc.comp[r][c] += c_zero * T.comp[r][c];
                    const int idx_mnu = NNM(n, idxh, nu, parity);
                    const int idx_pmu_mnu = NNM(n, idx_pmu, nu, !parity);
                    load_su3(g, u, 2 * nu + parity, idx_pmu_mnu, &A);
                    load_su3(g, u, 2 * mu + !parity, idx_mnu, &B);
                    load_su3(g, u, 2 * nu + !parity, idx_mnu, &C);
                    dag(&A, &A);
                    dag(&B, &B);
                    mul(&A, &B, &T);
                    mul(&T, &C, &T);
                    or_rebuild3row(&T);
                    for (int r = 0; r < 3; r++)
                        for (int c = 0; c < 3; c++) acc.comp[r][c] += c_zero * T.comp[r][c];
                }
                store_gl3(g, stap, dir_link, idxh, &acc, 3);
            }
    nn_free(&n);
}

/* traceless anti-Hermitian part, scaled (RHO_times_mat1_times_mat2_into_tamat3,
 * su3_utilities.h:1097-1152): QA = (rho/c_zero) TA(U * staple) */
static void ta_scaled(const or_su3 *X, double tmp, or_tamat *q)
{
    double tr = cimag(X->comp[0][0]) + cimag(X->comp[1][1]) + cimag(X->comp[2][2]);
    q->c01 = tmp * (0.5 * (X->comp[0][1] - conj(X->comp[1][0])));
    q->c02 = tmp * (0.5 * (X->comp[0][2] - conj(X->comp[2][0])));
    q->c12 = tmp * (0.5 * (X->comp[1][2] - conj(X->comp[2][1])));
    q->ic00 = tmp * (cimag(X->comp[0][0]) - OR_ONE_BY_THREE * tr);
    q->ic11 = tmp * (cimag(X->comp[1][1]) - OR_ONE_BY_THREE * tr);
}

void or_rho_times_conf_times_staples_ta(const or_geom *g, const void *u, const void *stap, void *qa,
                                        double rho, double c_zero)
{
    for (int link = 0; link < 8; link++)
        for (long i = 0; i < g->sizeh; i++) {
            or_su3 U, S, X;
            or_tamat q;
            load_su3(g, u, link, i, &U);
            load_gl3(g, stap, link, i, &S);
            mul(&U, &S, &X);
            ta_scaled(&X, rho / c_zero, &q);
            store_ta(g, qa, link, i, &q);
        }
}

/* exp_minus_QA_times_conf (stouting.c:107-165): rows 0,1 of exp(-QA) U */
void or_exp_minus_QA_times_conf(const or_geom *g, const void *u, const void *qa, void *uout)
{
    for (int link = 0; link < 8; link++)
        for (long i = 0; i < g->sizeh; i++) {
            or_tamat q;
            or_su3 E, U, O;
            load_ta(g, qa, link, i, &q);
            or_ch_exp_minus_QA(&q, &E);
            E.comp[2][0] = E.comp[2][1] = E.comp[2][2] = 0; /* only rows 0,1 of E are used */
            load_su3(g, u, link, i, &U);
            mul(&E, &U, &O);
            store_gl3(g, uout, link, i, &O, 2);
        }
}

/* stout_isotropic (stouting.c:74-97); staples/qa scratch may be NULL */
void or_stout_isotropic(const or_geom *g, const void *u, void *uout, double rho, double c_zero)
{
    void *stap = malloc(8 * or_su3_bytes(g));
    void *qa = malloc(8 * ta_bytes(g));
    or_calc_loc_staples(g, u, stap, c_zero);
    or_rho_times_conf_times_staples_ta(g, u, stap, qa, rho, c_zero);
    or_exp_minus_QA_times_conf(g, u, qa, uout);
    free(stap);
    free(qa);
}

/* ------------------------------------------------------------- stout force */
/* compute_loc_Lambda (stouting.c:171-512), Morningstar-Peardon eqs. 57-73 */
static void loc_lambda(const or_tamat *QA, const or_su3 *SP, const or_su3 *U, or_thmat *L)
{
    double c0 = or_det_i_times_QA(QA);
    double c1 = 0.5 * or_Tr_i_times_QA_sq(QA);
    double c0max = 2 * pow(c1 / 3, 1.5);
    or_dc f0, f1, f2, b10, b11, b12, b20, b21, b22;
    if (c1 < 4e-3) {
        f0 = (1 - c0 * c0 / 720) + (1.0 * I) * (-c0 * (1 - c1 * (1 - c1 / 42) / 20) / 6);
        f1 = (c0 * (1 - c1 * (1 - 3 * c1 / 112) / 15) / 24) + (1.0 * I) * (1 - c1 * (1 - c1 * (1 - c1 / 42) / 20) / 6 - c0 * c0 / 5040);
        f2 = (0.5 * (-1 + c1 * (1 - c1 * (1 - c1 / 56) / 30) / 12 + c0 * c0 / 20160)) + (1.0 * I) * (0.5 * (c0 * (1 - c1 * (1 - c1 / 48) / 21) / 60));
        b10 = 0.0 + (1.0 * I) * (c0 / 120 * (1 - c1 / 21));
        b11 = (-c0 / 360 * (1 - 3 * c1 / 56)) + (1.0 * I) * (-1.0 / 6 * (1 - c1 / 10 * (1.0 - c1 / 28)));
        b12 = (0.5 * (1.0 / 12 * (1 - 2 * c1 / 30 * (1 - 3 * c1 / 112)))) + (1.0 * I) * (0.5 * (-c0 / 1260 * (1 - c1 / 24)));
        b20 = (-c0 / 360) + (1.0 * I) * (-1.0 / 6 * (1 - c1 / 20 * (1 - c1 / 42)));
        b21 = (1.0 / 24 * (1 - c1 / 15 * (1 - 3 * c1 / 112))) + (1.0 * I) * (-c0 / 2520);
        b22 = (0.5 * c0 / 10080) + (1.0 * I) * (0.5 * (1.0 / 60 * (1 - c1 / 21 * (1 - c1 / 48))));
    } else {
        int sign = 1;
        if (c0 < 0) { sign = -1; c0 = -c0; }
        double eps = (c0max - c0) / c0max;
        double theta;
        if (eps < 0) theta = 0.0;
        else if (eps < 1e-3)
            theta = sqrt(2 * eps) * (1 + (1.0 / 12 + (3.0 / 160 + (5.0 / 896 + (35.0 / 18432 + 63.0 / 90112 * eps) * eps) * eps) * eps) * eps);
        else theta = acos(c0 / c0max);
        double u = sqrt(c1 / 3) * cos(theta / 3);
        double w = sqrt(c1) * sin(theta / 3);
        double u2 = u * u, w2 = w * w, u2mw2 = u2 - w2, w2p3u2 = w2 + 3 * u2, w2m3u2 = w2 - 3 * u2;
        double cu = cos(u), c2u = cos(2 * u);
        double su = sin(u), s2u = sin(2 * u);
        double cw = cos(w);
        double xi0w;
        if (fabs(w) < 0.05) {
            double temp0 = w * w, temp1 = 1 - temp0 / 42, temp2 = 1.0 - temp0 / 20 * temp1;
            xi0w = 1 - temp0 / 6 * temp2;
        } else xi0w = sin(w) / w;
        double xi1w;
        if (fabs(w) < 0.05) xi1w = -(1 - w2 * (1 - w2 * (1 - w2 / 54) / 28) / 10) / 3;
        else xi1w = cw / w2 - sin(w) / (w2 * w);
        double denom = 1 / (9 * u * u - w * w);
        f0 = (u2mw2 * c2u + cu * 8 * u2 * cw + 2 * su * u * w2p3u2 * xi0w) +
             (1.0 * I) * (u2mw2 * s2u + -su * 8 * u2 * cw + cu * 2 * u * w2p3u2 * xi0w);
        f0 *= denom;
        f1 = (2 * u * c2u + -cu * 2 * u * cw + -su * w2m3u2 * xi0w) + (1.0 * I) * (2 * u * s2u + su * 2 * u * cos(w) + -cu * w2m3u2 * xi0w);
        f1 *= denom;
        f2 = (c2u + -cu * cw + -3 * su * u * xi0w) + (1.0 * I) * (s2u + su * cw + -cu * 3 * u * xi0w);
        f2 *= denom;
        or_dc r0_1 = (2 * c2u * u + s2u * (-2 * u2 + 2 * w2) + 2 * cu * u * (8 * cw + 3 * u2 * xi0w + w2 * xi0w) + su * (-8 * cw * u2 + 18 * u2 * xi0w + 2 * w2 * xi0w)) +
                     (1.0 * I) * (-8 * cw * (2 * su * u + cu * u2) + 2 * (s2u * u + c2u * u2 - c2u * w2) + 2 * (9 * cu * u2 - 3 * su * u * u2 + cu * w2 - su * u * w2) * xi0w);
        or_dc r1_1 = (2 * c2u - 4 * s2u * u + su * (2 * cw * u + 6 * u * xi0w) + cu * (-2 * cw + 3 * u2 * xi0w - w2 * xi0w)) +
                     (1.0 * I) * (2 * s2u + 4 * c2u * u + 2 * cw * (su + cu * u) + (6 * cu * u - 3 * su * u2 + su * w2) * xi0w);
        or_dc r2_1 = (-2 * s2u + cw * su - 3 * (su + cu * u) * xi0w) + (1.0 * I) * (2 * c2u + cu * cw + (-3 * cu + 3 * su * u) * xi0w);
        or_dc r0_2 = (-2 * c2u + 2 * cw * su * u + 2 * su * u * xi0w - 8 * cu * u2 * xi0w + 6 * su * u * u2 * xi1w) +
                     (1.0 * I) * (2 * (-s2u + 4 * su * u2 * xi0w + cu * u * (cw + xi0w + 3 * u2 * xi1w)));
        or_dc r1_2 = (2 * cu * u * xi0w + su * (-cw - xi0w + 3 * u2 * xi1w)) + (1.0 * I) * (-2 * su * u * xi0w - cu * (cw + xi0w - 3 * u2 * xi1w));
        or_dc r2_2 = (cu * xi0w - 3 * su * u * xi1w) + (1.0 * I) * (-(su * xi0w) - 3 * cu * u * xi1w);
        b10 = 0.5 * denom * denom * (2.0 * u * r0_1 + (3.0 * u * u - w * w) * r0_2 - 2.0 * (15.0 * u * u + w * w) * f0);
        b11 = 0.5 * denom * denom * (2.0 * u * r1_1 + (3.0 * u * u - w * w) * r1_2 - 2.0 * (15.0 * u * u + w * w) * f1);
        b12 = 0.5 * denom * denom * (2.0 * u * r2_1 + (3.0 * u * u - w * w) * r2_2 - 2.0 * (15.0 * u * u + w * w) * f2);
        b20 = 0.5 * denom * denom * (r0_1 - 3.0 * u * r0_2 - 24.0 * u * f0);
        b21 = 0.5 * denom * denom * (r1_1 - 3.0 * u * r1_2 - 24.0 * u * f1);
        b22 = 0.5 * denom * denom * (r2_1 - 3.0 * u * r2_2 - 24.0 * u * f2);
        if (sign == -1) {
            b10 = +conj(b10); b11 = -conj(b11); b12 = +conj(b12);
            b20 = -conj(b20); b21 = +conj(b21); b22 = -conj(b22);
            f0 = conj(f0); f1 = -conj(f1); f2 = conj(f2);
        }
    }
    /* B1 = b10 + b11 Q + b12 Q^2, B2 likewise (eq. 69); Q = i QA */
    or_su3 B1, B2, Q, Q2, T, USP, G;
    or_itamat_poly(b10, b11, b12, QA, &B1);
    or_itamat_poly(b20, b21, b22, QA, &B2);
    or_itamat_poly(0.0, 1.0, 0.0, QA, &Q);
    or_itamat_poly(0.0, 0.0, 1.0, QA, &Q2);
    /* tr_k = Tr(Sigma' B_k U) */
    or_dc tr1 = 0, tr2 = 0;
    mul(&B1, U, &T);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) tr1 += SP->comp[a][b] * T.comp[b][a];
    mul(&B2, U, &T);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) tr2 += SP->comp[a][b] * T.comp[b][a];
    /* Gamma = tr1 Q + tr2 Q^2 + f1 U Sigma' + f2 (Q U Sigma' + U Sigma' Q) */
    mul(U, SP, &USP);
    or_su3 QU, UQ;
    mul(&Q, &USP, &QU);
    mul(&USP, &Q, &UQ);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            G.comp[a][b] = tr1 * Q.comp[a][b] + tr2 * Q2.comp[a][b] + f1 * USP.comp[a][b] + f2 * (QU.comp[a][b] + UQ.comp[a][b]);
    /* Lambda = (Gamma + Gamma^+)/2 - Tr(...)/6 */
    L->rc00 = (2 * creal(G.comp[0][0]) - creal(G.comp[1][1]) - creal(G.comp[2][2])) * OR_ONE_BY_THREE;
    L->rc11 = (2 * creal(G.comp[1][1]) - creal(G.comp[0][0]) - creal(G.comp[2][2])) * OR_ONE_BY_THREE;
    L->c01 = (G.comp[0][1] + conj(G.comp[1][0])) * 0.5;
    L->c02 = (G.comp[0][2] + conj(G.comp[2][0])) * 0.5;
    L->c12 = (G.comp[1][2] + conj(G.comp[2][1])) * 0.5;
    (void)f0;
}

void or_compute_lambda(const or_geom *g, void *L, const void *SP, const void *U, const void *QA)
{
    for (int link = 0; link < 8; link++)
        for (long i = 0; i < g->sizeh; i++) {
            or_tamat q;
            or_su3 sp, u;
            or_thmat l;
            load_ta(g, QA, link, i, &q);
            load_gl3(g, SP, link, i, &sp);
            load_su3(g, U, link, i, &u);
            loc_lambda(&q, &sp, &u, &l);
            store_th(g, L, link, i, &l);
        }
}

/* compute_sigma (stouting.c:1175-1302): Sigma = Sigma' exp(iQ) + i rho [
 *   sum_nu  A B C (L_mu(x) - L_nu(x)) + L_nu(x+mu) A B C - A B L_mu(x+nu) C        (right)
 *         + A' B' (L_nu(x-nu) - L_mu(x-nu)) C' + A' B' C' L_mu(x) - A' L_nu(x+mu-nu) B' C' ]  (left)
 * with A = U_nu(x+mu), B = U_mu(x+nu)^+, C = U_nu(x)^+,
 *      A' = U_nu(x+mu-nu)^+, B' = U_mu(x-nu)^+, C' = U_nu(x-nu). */
void or_compute_sigma(const or_geom *g, const void *L, const void *U, void *S, const void *QA, double rho)
{
    nn_t n = nn_make(g);
    const size_t sb = 8 * or_su3_bytes(g);
    void *Sin = malloc(sb);
    memcpy(Sin, S, sb);
    const or_dc irho = rho * I;
    for (long idxh = 0; idxh < g->sizeh; idxh++)
        for (int parity = 0; parity < 2; parity++)
            for (int mu = 0; mu < 4; mu++) {
                const int dir_link = 2 * mu + parity;
                /* PIECE1: Sigma' exp(iQ) with exp row 2 rebuilt */
                or_tamat q;
                or_su3 E, sp, res;
                load_ta(g, QA, dir_link, idxh, &q);
                or_ch_exp_minus_QA(&q, &E);
                or_rebuild3row(&E);
                load_gl3(g, Sin, dir_link, idxh, &sp);
                mul(&sp, &E, &res);
                or_thmat lt;
                or_su3 LD;
                load_th(g, L, dir_link, idxh, &lt);
                th_full(&lt, &LD);
                for (int iter = 0; iter < 3; iter++) {
                    static const int perp[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
                    const int nu = perp[mu][iter];
                    const int idx_pmu = NNP(n, idxh, mu, parity);
                    const int idx_pnu = NNP(n, idxh, nu, parity);
                    or_su3 A, B, C, ABC, T, X, LE, LF, LG;
                    load_su3(g, U, 2 * nu + !parity, idx_pmu, &A);
                    load_su3(g, U, 2 * mu + !parity, idx_pnu, &B);
                    dag(&B, &B);
                    load_su3(g, U, 2 * nu + parity, idxh, &C);
                    dag(&C, &C);
                    mul(&A, &B, &T);
                    mul(&T, &C, &ABC);
                    load_th(g, L, 2 * nu + parity, idxh, &lt);
                    th_full(&lt, &LE);
                    load_th(g, L, 2 * nu + !parity, idx_pmu, &lt);
                    th_full(&lt, &LF);
                    load_th(g, L, 2 * mu + !parity, idx_pnu, &lt);
                    th_full(&lt, &LG);
                    or_su3 DmE;
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) DmE.comp[a][b] = LD.comp[a][b] - LE.comp[a][b];
                    mul(&ABC, &DmE, &X);
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) res.comp[a][b] += irho * X.comp[a][b];
                    mul(&LF, &ABC, &X);
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) res.comp[a][b] += irho * X.comp[a][b];
                    mul(&T, &LG, &X); /* A B G */
                    mul(&X, &C, &X);
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) res.comp[a][b] -= irho * X.comp[a][b];
                    /* left */
                    const int idx_mnu = NNM(n, idxh, nu, parity);
                    const int idx_pmu_mnu = NNM(n, idx_pmu, nu, !parity);
                    load_su3(g, U, 2 * nu + parity, idx_pmu_mnu, &A);
                    dag(&A, &A);
                    load_su3(g, U, 2 * mu + !parity, idx_mnu, &B);
                    dag(&B, &B);
                    load_su3(g, U, 2 * nu + !parity, idx_mnu, &C);
                    or_su3 AB, LEl, LGl, LFl;
                    mul(&A, &B, &AB);
                    load_th(g, L, 2 * mu + !parity, idx_mnu, &lt);
                    th_full(&lt, &LEl);
                    load_th(g, L, 2 * nu + !parity, idx_mnu, &lt);
                    th_full(&lt, &LGl);
                    load_th(g, L, 2 * nu + parity, idx_pmu_mnu, &lt);
                    th_full(&lt, &LFl);
                    or_su3 GmE;
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) GmE.comp[a][b] = LGl.comp[a][b] - LEl.comp[a][b];
                    mul(&AB, &GmE, &X);
                    mul(&X, &C, &X);
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) res.comp[a][b] += irho * X.comp[a][b];
                    mul(&AB, &C, &X);
                    mul(&X, &LD, &X);
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) res.comp[a][b] += irho * X.comp[a][b];
                    mul(&A, &LFl, &X);
                    mul(&X, &B, &X);
                    mul(&X, &C, &X);
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) res.comp[a][b] -= irho * X.comp[a][b];
                }
                store_gl3(g, S, dir_link, idxh, &res, 3);
            }
    free(Sin);
    nn_free(&n);
}

/* compute_sigma_from_sigma_prime_backinto_sigma_prime (fermion_force.c:52-162) */
void or_sigma_prime_to_sigma(const or_geom *g, void *Sigma, void *Lambda, void *QA, const void *U, double rho,
                             double c_zero)
{
    void *tmp = malloc(8 * or_su3_bytes(g));
    or_calc_loc_staples(g, U, tmp, c_zero);
    or_rho_times_conf_times_staples_ta(g, U, tmp, QA, rho, c_zero);
    or_compute_lambda(g, Lambda, Sigma, U, QA);
    or_compute_sigma(g, Lambda, U, Sigma, QA, rho);
    free(tmp);
}

/* ------------------------------------------------------------ unitarize */
/* unitarize_conf (su3_utilities.c:17-40, loc_unitarize_conf su3_utilities.h:68-113) */
void or_unitarize_conf(const or_geom *g, void *u)
{
    for (int link = 0; link < 8; link++)
        for (long i = 0; i < g->sizeh; i++) {
            or_dc A00 = upl(g, u, link, 0, 0)[i], A01 = upl(g, u, link, 0, 1)[i], A02 = upl(g, u, link, 0, 2)[i];
            or_dc A10 = upl(g, u, link, 1, 0)[i], A11 = upl(g, u, link, 1, 1)[i], A12 = upl(g, u, link, 1, 2)[i];
            double NORM = creal(A00) * creal(A00) + cimag(A00) * cimag(A00) + creal(A01) * creal(A01) +
                          cimag(A01) * cimag(A01) + creal(A02) * creal(A02) + cimag(A02) * cimag(A02);
            NORM = 1.0 / sqrt(NORM);
            A00 *= NORM; A01 *= NORM; A02 *= NORM;
            or_dc SCAL_PROD = conj(A00) * A10 + conj(A01) * A11 + conj(A02) * A12;
            A10 -= SCAL_PROD * A00; A11 -= SCAL_PROD * A01; A12 -= SCAL_PROD * A02;
            NORM = creal(A10) * creal(A10) + cimag(A10) * cimag(A10) + creal(A11) * creal(A11) +
                   cimag(A11) * cimag(A11) + creal(A12) * creal(A12) + cimag(A12) * cimag(A12);
            NORM = 1.0 / sqrt(NORM);
            A10 *= NORM; A11 *= NORM; A12 *= NORM;
            upl(g, u, link, 0, 0)[i] = A00; upl(g, u, link, 0, 1)[i] = A01; upl(g, u, link, 0, 2)[i] = A02;
            upl(g, u, link, 1, 0)[i] = A10; upl(g, u, link, 1, 1)[i] = A11; upl(g, u, link, 1, 2)[i] = A12;
        }
}

/* ------------------------------------------------------ fermion force parts */
/* direct_product_of_fermions_into_auxmat (fermion_force_utilities.c:30-93):
 * even x: aux[2mu](x) += a * h(x+mu) (x) s(x)^+ ; odd x: aux[2mu+1](x) += -a * s(x+mu) (x) h(x)^+,
 * with s = x_k (even field) and h = Doe x_k (odd field), both vec3 of one parity. */
void or_direct_product_into_auxmat(const or_geom *g, const void *s, const void *h, void *aux, double a)
{
    nn_t n = nn_make(g);
    const or_dc *S[3], *Hh[3];
    for (int c = 0; c < 3; c++) {
        S[c] = (const or_dc *)((const char *)s + (size_t)c * 16 * g->sizeh);
        Hh[c] = (const or_dc *)((const char *)h + (size_t)c * 16 * g->sizeh);
    }
    for (int parity = 0; parity < 2; parity++)
        for (long idxh = 0; idxh < g->sizeh; idxh++)
            for (int mu = 0; mu < 4; mu++) {
                const int ip = NNP(n, idxh, mu, parity);
                const int dir = 2 * mu + parity;
                const or_dc *const *L = parity == 0 ? Hh : S;
                const or_dc *const *R = parity == 0 ? S : Hh;
                const double f = parity == 0 ? a : -a;
                or_dc r[3], l[3];
                for (int c = 0; c < 3; c++) {
                    r[c] = f * conj(R[c][idxh]);
                    l[c] = L[c][ip];
                }
                for (int a1 = 0; a1 < 3; a1++)
                    for (int b1 = 0; b1 < 3; b1++) upl(g, aux, dir, a1, b1)[idxh] += l[a1] * r[b1];
            }
    nn_free(&n);
}

/* multiply_backfield_times_force (fermion_force_utilities.c:123-153):
 * out[dir] += aux[dir] * e^{i phase[dir]} */
void or_multiply_backfield_times_force(const or_geom *g, const double *phases, const void *aux, void *out)
{
    for (int dir = 0; dir < 8; dir++)
        for (long i = 0; i < g->sizeh; i++) {
            double arg = phases[(size_t)dir * g->sizeh + i];
            or_dc phase = cos(arg) + I * sin(arg);
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) upl(g, out, dir, r, c)[i] += upl(g, aux, dir, r, c)[i] * phase;
        }
}

/* multiply_conf_times_force_and_take_ta_nophase (fermion_force_utilities.c:95-121,
 * mat1_times_auxmat_into_tamat_nophase .h:106-151): ipdot -= TA(U * aux) */
void or_conf_times_force_ta(const or_geom *g, const void *u, const void *aux, void *ipdot)
{
    for (int dir = 0; dir < 8; dir++)
        for (long i = 0; i < g->sizeh; i++) {
            or_su3 U, A, X;
            or_tamat q, t;
            load_su3(g, u, dir, i, &U);
            load_gl3(g, aux, dir, i, &A);
            mul(&U, &A, &X);
            ta_scaled(&X, 1.0, &t);
            load_ta(g, ipdot, dir, i, &q);
            q.c01 -= t.c01; q.c02 -= t.c02; q.c12 -= t.c12; q.ic00 -= t.ic00; q.ic11 -= t.ic11;
            store_ta(g, ipdot, dir, i, &q);
        }
}
