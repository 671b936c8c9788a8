/*
 * CPU oracle, single-site SU(3) helpers -- TEST INFRASTRUCTURE ONLY.
 * Restatements of OpenAcc/single_types.h and OpenAcc/cayley_hamilton.h.
 */
#ifndef OSTP_ORACLE_SU3_H
#define OSTP_ORACLE_SU3_H

#include "oracle.h"
#include <math.h>

typedef struct { or_dc comp[3][3]; } or_su3;
/* traceless anti-Hermitian matrix: upper off-diagonals and Im of diagonal 0,1
 * (single_tamat, single_types.h:28-34) */
typedef struct { or_dc c01, c02, c12; double ic00, ic11; } or_tamat;
/* traceless Hermitian matrix (single_thmat, single_types.h:35-41) */
typedef struct { or_dc c01, c02, c12; double rc00, rc11; } or_thmat;

/* rebuild3row, single_types.h:45-50 */
static inline void or_rebuild3row(or_su3 *m)
{
    m->comp[2][0] = conj(m->comp[0][1] * m->comp[1][2] - m->comp[0][2] * m->comp[1][1]);
    m->comp[2][1] = conj(m->comp[0][2] * m->comp[1][0] - m->comp[0][0] * m->comp[1][2]);
    m->comp[2][2] = conj(m->comp[0][0] * m->comp[1][1] - m->comp[0][1] * m->comp[1][0]);
}

/* det_i_times_QA / Tr_i_times_QA_sq on a single tamat (cayley_hamilton.h:24-50) */
static inline double or_det_i_times_QA(const or_tamat *QA)
{
    double ic22 = -QA->ic00 - QA->ic11;
    return -creal(QA->ic00 * QA->ic11 * ic22
                  + 2 * cimag(QA->c01 * QA->c12 * conj(QA->c02))
                  - QA->ic00 * QA->c12 * conj(QA->c12)
                  - QA->ic11 * QA->c02 * conj(QA->c02)
                  - QA->c01 * conj(QA->c01) * ic22);
}
static inline double or_Tr_i_times_QA_sq(const or_tamat *QA)
{
    return 2 * creal(QA->ic00 * QA->ic00 + QA->ic11 * QA->ic11 + QA->ic00 * QA->ic11 +
                     QA->c01 * conj(QA->c01) + QA->c02 * conj(QA->c02) + QA->c12 * conj(QA->c12));
}

/* f0,f1,f2 of exp(iQ), Q = i*QA (cayley_hamilton.h:66-147; Morningstar-Peardon
 * hep-lat/0311018 eqs. 14-34 with the small-c1 series taken from NISSA) */
static inline void or_ch_coefficients(const or_tamat *QA, or_dc *pf0, or_dc *pf1, or_dc *pf2)
{
    double c0 = or_det_i_times_QA(QA);
    double c1 = 0.5 * or_Tr_i_times_QA_sq(QA);
    double c0max = 2 * pow(c1 / 3, 1.5);
    or_dc f0, f1, f2;
    if (c1 < 4.0e-3) {
        f0 = (1 - c0 * c0 / 720) + (1.0 * I) * (-c0 * (1 - c1 * (1 - c1 / 42) / 20) / 6);
        f1 = (c0 * (1 - c1 * (1 - 3 * c1 / 112) / 15) / 24) +
             (1.0 * I) * (1 - c1 * (1 - c1 * (1 - c1 / 42) / 20) / 6 - c0 * c0 / 5040);
        f2 = (0.5 * (-1 + c1 * (1 - c1 * (1 - c1 / 56) / 30) / 12 + c0 * c0 / 20160)) +
             (1.0 * I) * (0.5 * (c0 * (1 - c1 * (1 - c1 / 48) / 21) / 60));
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
        double denom = 1 / (9 * u * u - w * w);
        f0 = (u2mw2 * c2u + cu * 8 * u2 * cw + 2 * su * u * w2p3u2 * xi0w) +
             (1.0 * I) * (u2mw2 * s2u + -su * 8 * u2 * cw + cu * 2 * u * w2p3u2 * xi0w);
        f0 *= denom;
        f1 = (2 * u * c2u + -cu * 2 * u * cw + -su * w2m3u2 * xi0w) +
             (1.0 * I) * (2 * u * s2u + su * 2 * u * cos(w) + -cu * w2m3u2 * xi0w);
        f1 *= denom;
        f2 = (c2u + -cu * cw + -3 * su * u * xi0w) + (1.0 * I) * (s2u + su * cw + -cu * 3 * u * xi0w);
        f2 *= denom;
        if (sign == -1) { f0 = conj(f0); f1 = -conj(f1); f2 = conj(f2); }
    }
    *pf0 = f0; *pf1 = f1; *pf2 = f2;
}

/* Itamat_2ndDeg_poly_no3rdrow (cayley_hamilton.h:186-222): rows 0,1 of
 * f0 + f1 Q + f2 Q^2 */
static inline void or_itamat_poly_no3rdrow(or_dc f0, or_dc f1, or_dc f2, const or_tamat *QA, or_su3 *out)
{
    out->comp[0][0] = f0 - f1 * QA->ic00 + f2 * (QA->ic00 * QA->ic00 + QA->c01 * conj(QA->c01) + QA->c02 * conj(QA->c02));
    out->comp[0][1] = (f1 * I) * QA->c01 + f2 * (QA->c02 * conj(QA->c12) + (-1.0 * I) * QA->c01 * (QA->ic00 + QA->ic11));
    out->comp[0][2] = (f1 * I) * QA->c02 + f2 * (-QA->c01 * QA->c12 + (1.0 * I) * QA->c02 * QA->ic11);
    out->comp[1][0] = (-f1 * I) * conj(QA->c01) + f2 * (QA->c12 * conj(QA->c02) + (1.0 * I) * conj(QA->c01) * (QA->ic00 + QA->ic11));
    out->comp[1][1] = f0 - f1 * QA->ic11 + f2 * (QA->ic11 * QA->ic11 + QA->c01 * conj(QA->c01) + QA->c12 * conj(QA->c12));
    out->comp[1][2] = (f1 * I) * QA->c12 + f2 * ((1.0 * I) * QA->ic00 * QA->c12 + QA->c02 * conj(QA->c01));
}

/* Itamat_2ndDeg_poly (cayley_hamilton.h:225-243): adds row 2 */
static inline void or_itamat_poly(or_dc f0, or_dc f1, or_dc f2, const or_tamat *QA, or_su3 *out)
{
    or_itamat_poly_no3rdrow(f0, f1, f2, QA, out);
    out->comp[2][0] = (f1 * I) * (-conj(QA->c02)) - f2 * conj(QA->c01 * QA->c12 - QA->c02 * QA->ic11 * I);
    out->comp[2][1] = (f1 * I) * (-conj(QA->c12)) - f2 * (-conj(QA->c02) * QA->c01 + conj(QA->c12) * QA->ic00 * I);
    out->comp[2][2] = f0 + f1 * (QA->ic00 + QA->ic11) +
                      f2 * (QA->c02 * conj(QA->c02) + QA->c12 * conj(QA->c12) + (QA->ic00 + QA->ic11) * (QA->ic00 + QA->ic11));
}

/* CH_exponential_antihermitian_nissalike: rows 0,1 of exp(-QA); row 2 rebuilt by caller */
static inline void or_ch_exp_minus_QA(const or_tamat *QA, or_su3 *out)
{
    or_dc f0, f1, f2;
    or_ch_coefficients(QA, &f0, &f1, &f2);
    or_itamat_poly_no3rdrow(f0, f1, f2, QA, out);
}

#endif
