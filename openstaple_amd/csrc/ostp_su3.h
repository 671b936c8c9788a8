// Single-site SU(3) algebra on the device (FP64): 3x3 complex matrices, the
// traceless anti-Hermitian (tamat) / Hermitian (thmat) compact forms of
// OpenAcc/single_types.h:24-41, and the Cayley-Hamilton exponential of
// OpenAcc/cayley_hamilton.h:52-243 (Morningstar-Peardon, hep-lat/0311018,
// Sec. III, with the small-c1 series used by MILC/NISSA).
#pragma once

#include "ostp_common.h"

namespace ostp {

struct Su3 {
    double2 c[3][3];
};
// QA = [[i ic00, c01, c02], [-c01*, i ic11, c12], [-c02*, -c12*, i ic22]], ic22 = -ic00-ic11
struct Tamat {
    double2 c01, c02, c12;
    double ic00, ic11;
};
// H = [[rc00, c01, c02], [c01*, rc11, c12], [c02*, c12*, rc22]], rc22 = -rc00-rc11
struct Thmat {
    double2 c01, c02, c12;
    double rc00, rc11;
};

__device__ __forceinline__ double2 cmul_nc(double2 a, double2 b) {
#pragma clang fp contract(off)
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 c_i(double2 a) { return make_double2(-a.y, a.x); }   // i*a
__device__ __forceinline__ double2 c_mi(double2 a) { return make_double2(a.y, -a.x); }  // -i*a
__device__ __forceinline__ double2 creal2(double r) { return make_double2(r, 0.0); }
__device__ __forceinline__ double cnorm2(double2 a) { return a.x * a.x + a.y * a.y; }

// rebuild3row (single_types.h:45-50): row2 = conj(row0 x row1)
__device__ __forceinline__ void rebuild3row(Su3& m) {
    m.c[2][0] = cconj(csub(cmul(m.c[0][1], m.c[1][2]), cmul(m.c[0][2], m.c[1][1])));
    m.c[2][1] = cconj(csub(cmul(m.c[0][2], m.c[1][0]), cmul(m.c[0][0], m.c[1][2])));
    m.c[2][2] = cconj(csub(cmul(m.c[0][0], m.c[1][1]), cmul(m.c[0][1], m.c[1][0])));
}

// det(i QA) and Tr((i QA)^2)  (cayley_hamilton.h:24-50)
__device__ __forceinline__ double det_i_times_qa(const Tamat& q) {
    const double ic22 = -q.ic00 - q.ic11;
    const double2 t = cmul(cmul(q.c01, q.c12), cconj(q.c02));
    return -(q.ic00 * q.ic11 * ic22 + 2 * t.y - q.ic00 * cnorm2(q.c12) - q.ic11 * cnorm2(q.c02) - cnorm2(q.c01) * ic22);
}
__device__ __forceinline__ double tr_i_times_qa_sq(const Tamat& q) {
    return 2 * (q.ic00 * q.ic00 + q.ic11 * q.ic11 + q.ic00 * q.ic11 + cnorm2(q.c01) + cnorm2(q.c02) + cnorm2(q.c12));
}

// f0, f1, f2 with exp(iQ) = f0 + f1 Q + f2 Q^2, Q = i QA  (cayley_hamilton.h:66-147)
__device__ __forceinline__ void ch_coefficients(const Tamat& q, double2& f0, double2& f1, double2& f2) {
    double c0 = det_i_times_qa(q);
    const double c1 = 0.5 * tr_i_times_qa_sq(q);
    const double c0max = 2 * pow(c1 / 3, 1.5);
    if (c1 < 4.0e-3) {
        f0 = make_double2(1 - c0 * c0 / 720, -c0 * (1 - c1 * (1 - c1 / 42) / 20) / 6);
        f1 = make_double2(c0 * (1 - c1 * (1 - 3 * c1 / 112) / 15) / 24,
                          1 - c1 * (1 - c1 * (1 - c1 / 42) / 20) / 6 - c0 * c0 / 5040);
        f2 = make_double2(0.5 * (-1 + c1 * (1 - c1 * (1 - c1 / 56) / 30) / 12 + c0 * c0 / 20160),
                          0.5 * (c0 * (1 - c1 * (1 - c1 / 48) / 21) / 60));
        return;
    }
    int sign = 1;
    if (c0 < 0) {
        sign = -1;
        c0 = -c0;
    }
    const double eps = (c0max - c0) / c0max;
    double theta;
    if (eps < 0) theta = 0.0;
    else if (eps < 1e-3)
        theta = sqrt(2 * eps) *
                (1 + (1.0 / 12 + (3.0 / 160 + (5.0 / 896 + (35.0 / 18432 + 63.0 / 90112 * eps) * eps) * eps) * eps) * eps);
    else theta = acos(c0 / c0max);
    const double u = sqrt(c1 / 3) * cos(theta / 3);
    const double w = sqrt(c1) * sin(theta / 3);
    const double u2 = u * u, w2 = w * w, u2mw2 = u2 - w2, w2p3u2 = w2 + 3 * u2, w2m3u2 = w2 - 3 * u2;
    const double cu = cos(u), c2u = cos(2 * u);
    const double su = sin(u), s2u = sin(2 * u);
    const double cw = cos(w);
    double xi0w;
    if (fabs(w) < 0.05) {
        const double temp0 = w * w, temp1 = 1 - temp0 / 42, temp2 = 1.0 - temp0 / 20 * temp1;
        xi0w = 1 - temp0 / 6 * temp2;
    } else {
        xi0w = sin(w) / w;
    }
    const double denom = 1 / (9 * u * u - w * w);
    f0 = make_double2((u2mw2 * c2u + cu * 8 * u2 * cw + 2 * su * u * w2p3u2 * xi0w) * denom,
                      (u2mw2 * s2u + -su * 8 * u2 * cw + cu * 2 * u * w2p3u2 * xi0w) * denom);
    f1 = make_double2((2 * u * c2u + -cu * 2 * u * cw + -su * w2m3u2 * xi0w) * denom,
                      (2 * u * s2u + su * 2 * u * cw + -cu * w2m3u2 * xi0w) * denom);
    f2 = make_double2((c2u + -cu * cw + -3 * su * u * xi0w) * denom, (s2u + su * cw + -cu * 3 * u * xi0w) * denom);
    if (sign == -1) {
        f0 = cconj(f0);
        f1 = make_double2(-f1.x, f1.y);
        f2 = cconj(f2);
    }
}

// rows 0,1 of f0 + f1 Q + f2 Q^2 (Itamat_2ndDeg_poly_no3rdrow, cayley_hamilton.h:186-222)
__device__ __forceinline__ void itamat_poly_no3rdrow(double2 f0, double2 f1, double2 f2, const Tamat& q, Su3& o) {
    const double2 if1 = c_i(f1);
    const double s01 = q.ic00 + q.ic11;
    o.c[0][0] = cadd(csub(f0, cscale(f1, q.ic00)),
                     cscale(f2, q.ic00 * q.ic00 + cnorm2(q.c01) + cnorm2(q.c02)));
    o.c[0][1] = cadd(cmul(if1, q.c01), cmul(f2, cadd(cmul(q.c02, cconj(q.c12)), cscale(c_mi(q.c01), s01))));
    o.c[0][2] = cadd(cmul(if1, q.c02), cmul(f2, cadd(make_double2(-cmul(q.c01, q.c12).x, -cmul(q.c01, q.c12).y),
                                                      cscale(c_i(q.c02), q.ic11))));
    o.c[1][0] = cadd(cmul(make_double2(-if1.x, -if1.y), cconj(q.c01)),
                     cmul(f2, cadd(cmul(q.c12, cconj(q.c02)), cscale(c_i(cconj(q.c01)), s01))));
    o.c[1][1] = cadd(csub(f0, cscale(f1, q.ic11)),
                     cscale(f2, q.ic11 * q.ic11 + cnorm2(q.c01) + cnorm2(q.c12)));
    o.c[1][2] = cadd(cmul(if1, q.c12), cmul(f2, cadd(cscale(c_i(q.c12), q.ic00), cmul(q.c02, cconj(q.c01)))));
}

// row 2 (Itamat_2ndDeg_poly, cayley_hamilton.h:225-243)
__device__ __forceinline__ void itamat_poly_row2(double2 f0, double2 f1, double2 f2, const Tamat& q, Su3& o) {
    const double2 if1 = c_i(f1);
    const double s01 = q.ic00 + q.ic11;
    o.c[2][0] = csub(cmul(if1, make_double2(-q.c02.x, q.c02.y)),
                     cmul(f2, cconj(csub(cmul(q.c01, q.c12), cscale(c_i(q.c02), q.ic11)))));
    o.c[2][1] = csub(cmul(if1, make_double2(-q.c12.x, q.c12.y)),
                     cmul(f2, cadd(make_double2(-cmul(cconj(q.c02), q.c01).x, -cmul(cconj(q.c02), q.c01).y),
                                   cscale(c_i(cconj(q.c12)), q.ic00))));
    o.c[2][2] = cadd(cadd(f0, cscale(f1, s01)), cscale(f2, cnorm2(q.c02) + cnorm2(q.c12) + s01 * s01));
}

// rows 0,1 of exp(-QA); row 2 by rebuild3row or itamat_poly_row2
__device__ __forceinline__ void ch_exp_minus_qa(const Tamat& q, Su3& o) {
    double2 f0, f1, f2;
    ch_coefficients(q, f0, f1, f2);
    itamat_poly_no3rdrow(f0, f1, f2, q, o);
}

// ---------------------------------------------------------- 3x3 algebra
__device__ __forceinline__ void su3_zero(Su3& m) {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) m.c[r][c] = make_double2(0.0, 0.0);
}
// o = a * b
__device__ __forceinline__ void su3_mul(const Su3& a, const Su3& b, Su3& o) {
    Su3 t;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double2 s = cmul(a.c[r][0], b.c[0][c]);
            s = cfma(a.c[r][1], b.c[1][c], s);
            s = cfma(a.c[r][2], b.c[2][c], s);
            t.c[r][c] = s;
        }
    o = t;
}
// o = a * b^+
__device__ __forceinline__ void su3_mul_dag(const Su3& a, const Su3& b, Su3& o) {
    Su3 t;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double2 s = cmul(a.c[r][0], cconj(b.c[c][0]));
            s = cfma(a.c[r][1], cconj(b.c[c][1]), s);
            s = cfma(a.c[r][2], cconj(b.c[c][2]), s);
            t.c[r][c] = s;
        }
    o = t;
}
// o = a^+ * b
__device__ __forceinline__ void su3_dag_mul(const Su3& a, const Su3& b, Su3& o) {
    Su3 t;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double2 s = cmulc(a.c[0][r], b.c[0][c]);
            s = cfmac(a.c[1][r], b.c[1][c], s);
            s = cfmac(a.c[2][r], b.c[2][c], s);
            t.c[r][c] = s;
        }
    o = t;
}
// acc += s * m  (complex scalar)
__device__ __forceinline__ void su3_axpy(double2 s, const Su3& m, Su3& acc) {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) acc.c[r][c] = cfma(s, m.c[r][c], acc.c[r][c]);
}
__device__ __forceinline__ void su3_scale_add(double s, const Su3& m, Su3& acc) {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            acc.c[r][c] = make_double2(fma(s, m.c[r][c].x, acc.c[r][c].x), fma(s, m.c[r][c].y, acc.c[r][c].y));
}
// full Hermitian matrix of a thmat
__device__ __forceinline__ void th_full(const Thmat& l, Su3& o) {
    o.c[0][0] = creal2(l.rc00);
    o.c[0][1] = l.c01;
    o.c[0][2] = l.c02;
    o.c[1][0] = cconj(l.c01);
    o.c[1][1] = creal2(l.rc11);
    o.c[1][2] = l.c12;
    o.c[2][0] = cconj(l.c02);
    o.c[2][1] = cconj(l.c12);
    o.c[2][2] = creal2(-l.rc00 - l.rc11);
}
// scaled traceless anti-Hermitian part (RHO_times_mat1_times_mat2_into_tamat3)
__device__ __forceinline__ void ta_part(const Su3& X, double s, Tamat& q) {
    const double tr = X.c[0][0].y + X.c[1][1].y + X.c[2][2].y;
    const double third = 0.33333333333333333333333;
    q.c01 = cscale(csub(X.c[0][1], cconj(X.c[1][0])), 0.5 * s);
    q.c02 = cscale(csub(X.c[0][2], cconj(X.c[2][0])), 0.5 * s);
    q.c12 = cscale(csub(X.c[1][2], cconj(X.c[2][1])), 0.5 * s);
    q.ic00 = s * (X.c[0][0].y - third * tr);
    q.ic11 = s * (X.c[1][1].y - third * tr);
}

// ---------------------------------------------- tamat_soa / thmat_soa access
// tamat / thmat soa: c01,c02,c12 complex planes then two real planes
struct TaPtr {
    double2* c;   // 3 planes
    double* d;    // 2 planes
};
__device__ __forceinline__ TaPtr ta_at(void* base, int link, int sizeh) {
    char* b = (char*)base + (size_t)link * 64 * sizeh;
    return TaPtr{(double2*)b, (double*)(b + (size_t)48 * sizeh)};
}
__device__ __forceinline__ void ta_load(const void* base, int link, int sizeh, int i, Tamat& q) {
    TaPtr p = ta_at(const_cast<void*>(base), link, sizeh);
    q.c01 = p.c[i];
    q.c02 = p.c[sizeh + i];
    q.c12 = p.c[2 * sizeh + i];
    q.ic00 = p.d[i];
    q.ic11 = p.d[sizeh + i];
}
__device__ __forceinline__ void ta_store(void* base, int link, int sizeh, int i, const Tamat& q) {
    TaPtr p = ta_at(base, link, sizeh);
    p.c[i] = q.c01;
    p.c[sizeh + i] = q.c02;
    p.c[2 * sizeh + i] = q.c12;
    p.d[i] = q.ic00;
    p.d[sizeh + i] = q.ic11;
}
__device__ __forceinline__ void th_load(const void* base, int link, int sizeh, int i, Thmat& l) {
    TaPtr p = ta_at(const_cast<void*>(base), link, sizeh);
    l.c01 = p.c[i];
    l.c02 = p.c[sizeh + i];
    l.c12 = p.c[2 * sizeh + i];
    l.rc00 = p.d[i];
    l.rc11 = p.d[sizeh + i];
}
__device__ __forceinline__ void th_store(void* base, int link, int sizeh, int i, const Thmat& l) {
    TaPtr p = ta_at(base, link, sizeh);
    p.c[i] = l.c01;
    p.c[sizeh + i] = l.c02;
    p.c[2 * sizeh + i] = l.c12;
    p.d[i] = l.rc00;
    p.d[sizeh + i] = l.rc11;
}

// thm -= (i fact) tam on c01, c02, c12; rc00/rc11 += fact ic00/ic11, rounded
// like the reference's complex product (no fused multiply-add)
__device__ __forceinline__ void mom_step(void* mom, int link, int sizeh, int i, const Tamat& q, double fact) {
#pragma clang fp contract(off)
    Thmat h;
    th_load(mom, link, sizeh, i, h);
    h.c01 = make_double2(h.c01.x + fact * q.c01.y, h.c01.y - fact * q.c01.x);
    h.c02 = make_double2(h.c02.x + fact * q.c02.y, h.c02.y - fact * q.c02.x);
    h.c12 = make_double2(h.c12.x + fact * q.c12.y, h.c12.y - fact * q.c12.x);
    h.rc00 = h.rc00 + fact * q.ic00;
    h.rc11 = h.rc11 + fact * q.ic11;
    th_store(mom, link, sizeh, i, h);
}

}  // namespace ostp
