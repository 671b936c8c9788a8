/*
 * CPU oracle, fermion part -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Each function restates the reference loop it cites, in the same order of
 * floating-point operations as the reference's gcc build (OpenACC pragmas are
 * ignored by gcc, so that build is this serial loop nest).  Built with
 * -O2 -ffp-contract=off so no fused multiply-add changes the rounding.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_TWOPI (2 * 3.14159265358979323846) /* backfield.c:11 acc_twopi */

/* ------------------------------------------------------------------ layout */

void or_geom_init(or_geom *g, int nd0, int nd1, int nd2, int nd3, int par_temp)
{
    g->nd[0] = nd0; g->nd[1] = nd1; g->nd[2] = nd2; g->nd[3] = nd3;
    g->sizeh = (long)nd0 * nd1 * nd2 * nd3 / 2;
    g->par_temp = par_temp;
}

size_t or_vec3_bytes(const or_geom *g) { return (size_t)48 * g->sizeh; }
size_t or_su3_bytes(const or_geom *g) { return (size_t)(144 + (g->par_temp ? 8 : 0)) * g->sizeh; }

static inline or_dc *vplane(const or_geom *g, const void *v, int c)
{
    return (or_dc *)((char *)v + (size_t)c * 16 * g->sizeh);
}
static inline or_fc *vplane_f(const or_geom *g, const void *v, int c)
{
    return (or_fc *)((char *)v + (size_t)c * 8 * g->sizeh);
}
/* link array `link` (0..7), row r, column c */
static inline or_dc *uplane(const or_geom *g, const void *u, int link, int r, int c)
{
    return (or_dc *)((char *)u + (size_t)link * or_su3_bytes(g) + (size_t)(3 * r + c) * 16 * g->sizeh);
}
static inline size_t su3_bytes_f(const or_geom *g) { return (size_t)(72 + (g->par_temp ? 4 : 0)) * g->sizeh; }
static inline or_fc *uplane_f(const or_geom *g, const void *u, int link, int r, int c)
{
    return (or_fc *)((char *)u + (size_t)link * su3_bytes_f(g) + (size_t)(3 * r + c) * 8 * g->sizeh);
}

/* --------------------------------------------------------------- geometry */

/* snum_acc, Mpi/geometry_multidev.h:219-229 (single rank: LNH == LOC) */
int or_snum(const or_geom *g, int d0, int d1, int d2, int d3)
{
    int vol1 = g->nd[0], vol2 = g->nd[1] * vol1, vol3 = g->nd[2] * vol2;
    int ris = d0 + d1 * vol1 + d2 * vol2 + d3 * vol3;
    return ris / 2;
}

/* compute_nnp_and_nnm_openacc, OpenAcc/geometry.c:14-100 (NRANKS_D* == 1) */
void or_nn_tables(const or_geom *g, int *nnp, int *nnm)
{
    const int nd0 = g->nd[0], nd1 = g->nd[1], nd2 = g->nd[2], nd3 = g->nd[3];
    for (int d3 = 0; d3 < nd3; d3++)
        for (int d2 = 0; d2 < nd2; d2++)
            for (int d1 = 0; d1 < nd1; d1++)
                for (int d0 = 0; d0 < nd0; d0++) {
                    int idxh = or_snum(g, d0, d1, d2, d3);
                    int parity = (d0 + d1 + d2 + d3) % 2;
                    int d0m = d0 - 1, d1m = d1 - 1, d2m = d2 - 1, d3m = d3 - 1;
                    d0m = d0m + (((d0m >> 31) & 0x1) * nd0);
                    d1m = d1m + (((d1m >> 31) & 0x1) * nd1);
                    d2m = d2m + (((d2m >> 31) & 0x1) * nd2);
                    d3m = d3m + (((d3m >> 31) & 0x1) * nd3);
                    int *m = nnm + (size_t)idxh * 8;
                    m[0 * 2 + parity] = or_snum(g, d0m, d1, d2, d3);
                    m[1 * 2 + parity] = or_snum(g, d0, d1m, d2, d3);
                    m[2 * 2 + parity] = or_snum(g, d0, d1, d2m, d3);
                    m[3 * 2 + parity] = or_snum(g, d0, d1, d2, d3m);
                    int d0p = d0 + 1, d1p = d1 + 1, d2p = d2 + 1, d3p = d3 + 1;
                    d0p *= (((d0p - nd0) >> 31) & 0x1);
                    d1p *= (((d1p - nd1) >> 31) & 0x1);
                    d2p *= (((d2p - nd2) >> 31) & 0x1);
                    d3p *= (((d3p - nd3) >> 31) & 0x1);
                    int *p = nnp + (size_t)idxh * 8;
                    p[0 * 2 + parity] = or_snum(g, d0p, d1, d2, d3);
                    p[1 * 2 + parity] = or_snum(g, d0, d1p, d2, d3);
                    p[2 * 2 + parity] = or_snum(g, d0, d1, d2p, d3);
                    p[3 * 2 + parity] = or_snum(g, d0, d1, d2, d3p);
                }
}

/* ----------------------------------------------------------------- phases */

static int ks_x(int x, int y, int z, int t) { (void)x; (void)y; (void)z; (void)t; return 1; }
static int ks_y(int x, int y, int z, int t) { (void)y; (void)z; (void)t; return 1 - (2 * (x & 0x1)); }
static int ks_z(int x, int y, int z, int t) { (void)z; (void)t; return 1 - (2 * ((x + y) & 0x1)); }
static int ks_t(int x, int y, int z, int t) { (void)t; return 1 - (2 * ((x + y + z) & 0x1)); }

/* calc_u1_phases = calc_u1_phases_unb_no2pi (backfield.c:20-160)
 *                + rebound_u1_phases (:163-170) + mult_u1_phases (:172-177),
 * KSphase* from backfield.h:27-41. Single rank (no origin shift). */
void or_calc_u1_phases(const or_geom *g, const int map[4], const double bf[6],
                       double im_chem_pot, double ferm_charge, double *phases)
{
    const double ex_quantum = bf[0], ey_quantum = bf[1], ez_quantum = bf[2];
    const double bx_quantum = bf[3], by_quantum = bf[4], bz_quantum = bf[5];
    const int xmap = map[0], ymap = map[1], zmap = map[2], tmap = map[3];
    const int gnt = g->nd[tmap];
    const long sizeh = g->sizeh;
    double chpotphase = im_chem_pot / gnt;
    int d[4];
    for (d[3] = 0; d[3] < g->nd[3]; d[3]++)
        for (d[2] = 0; d[2] < g->nd[2]; d[2]++)
            for (d[1] = 0; d[1] < g->nd[1]; d[1]++)
                for (d[0] = 0; d[0] < g->nd[0]; d[0]++) {
                    int idxh = or_snum(g, d[0], d[1], d[2], d[3]);
                    int x = d[xmap]; int tnx = g->nd[xmap];
                    int y = d[ymap]; int tny = g->nd[ymap];
                    int z = d[zmap]; int tnz = g->nd[zmap];
                    int t = d[tmap]; int tnt = g->nd[tmap];
                    int parity = (x + y + z + t) % 2;
                    int X = x + 1, Y = y + 1, Z = z + 1, T = t + 1;
                    double arg;

                    arg = (z - tnz / 2 + 1) * by_quantum / (tnz * tnx);
                    if (X == tnx) {
                        arg -= (y - tny / 2 + 1) * tnx * bz_quantum / (tnx * tny);
                        arg -= (t - tnt / 2 + 1) * tnx * ex_quantum / (tnx * tnt);
                    }
                    arg *= ferm_charge;
                    if (ks_x(x, y, z, t) == -1) arg += 0.5;
                    phases[(size_t)(xmap * 2 + parity) * sizeh + idxh] = arg;

                    arg = (x - tnx / 2 + 1) * bz_quantum / (tnx * tny);
                    if (Y == tny) {
                        arg -= (z - tnz / 2 + 1) * tny * bx_quantum / (tny * tnz);
                        arg -= (t - tnt / 2 + 1) * tny * ey_quantum / (tny * tnt);
                    }
                    arg *= ferm_charge;
                    if (ks_y(x, y, z, t) == -1) arg += 0.5;
                    phases[(size_t)(ymap * 2 + parity) * sizeh + idxh] = arg;

                    arg = (y - tny / 2 + 1) * bx_quantum / (tny * tnz);
                    if (Z == tnz) {
                        arg -= (t - tnt / 2 + 1) * tnz * ez_quantum / (tnz * tnt);
                        arg -= (x - tnx / 2 + 1) * tnz * by_quantum / (tnz * tnx);
                    }
                    arg *= ferm_charge;
                    if (ks_z(x, y, z, t) == -1) arg += 0.5;
                    phases[(size_t)(zmap * 2 + parity) * sizeh + idxh] = arg;

                    arg = (z - tnz / 2 + 1) * ez_quantum / (tnz * tnt);
                    arg += (y - tny / 2 + 1) * ey_quantum / (tny * tnt);
                    arg += (x - tnx / 2 + 1) * ex_quantum / (tnx * tnt);
                    arg *= ferm_charge;
                    if (ks_t(x, y, z, t) == -1) arg += 0.5;
                    arg += chpotphase * 0.5;
                    if (T == tnt) arg += 0.5;
                    phases[(size_t)(tmap * 2 + parity) * sizeh + idxh] = arg;
                }
    for (int dir = 0; dir < 8; dir++)
        for (long i = 0; i < sizeh; i++) {
            double *a = &phases[(size_t)dir * sizeh + i];
            while (*a > 0.5) *a -= 1.0;
            while (*a < -0.5) *a += 1.0;
        }
    for (int dir = 0; dir < 8; dir++)
        for (long i = 0; i < sizeh; i++) phases[(size_t)dir * sizeh + i] *= OR_TWOPI;
}

void or_phases_to_float(const or_geom *g, const double *phases, float *phases_f)
{
    for (long i = 0; i < 8 * g->sizeh; i++) phases_f[i] = (float)phases[i];
}

/* ----------------------------------------------------------------- Dslash */

typedef struct { or_dc c0, c1, c2; } or_vec3;

/* mat_vec_mul_arg, matvecmul.h:88-126 */
static inline or_vec3 mat_vec_mul_arg(const or_geom *g, const void *u, int link, int idx_mat,
                                      const void *in, int idx_vect, const double *phases)
{
    or_vec3 o;
    double arg = phases[(size_t)link * g->sizeh + idx_mat];
    or_dc phase = cos(arg) + I * sin(arg);
    or_dc vec0 = vplane(g, in, 0)[idx_vect] * phase;
    or_dc vec1 = vplane(g, in, 1)[idx_vect] * phase;
    or_dc vec2 = vplane(g, in, 2)[idx_vect] * phase;
    or_dc mat00 = uplane(g, u, link, 0, 0)[idx_mat];
    or_dc mat01 = uplane(g, u, link, 0, 1)[idx_mat];
    or_dc mat02 = uplane(g, u, link, 0, 2)[idx_mat];
    or_dc mat10 = uplane(g, u, link, 1, 0)[idx_mat];
    or_dc mat11 = uplane(g, u, link, 1, 1)[idx_mat];
    or_dc mat12 = uplane(g, u, link, 1, 2)[idx_mat];
    or_dc mat20 = conj((mat01 * mat12) - (mat02 * mat11));
    or_dc mat21 = conj((mat02 * mat10) - (mat00 * mat12));
    or_dc mat22 = conj((mat00 * mat11) - (mat01 * mat10));
    o.c0 = (mat00 * vec0) + (mat01 * vec1) + (mat02 * vec2);
    o.c1 = (mat10 * vec0) + (mat11 * vec1) + (mat12 * vec2);
    o.c2 = (mat20 * vec0) + (mat21 * vec1) + (mat22 * vec2);
    return o;
}

/* conjmat_vec_mul_arg, matvecmul.h:129-172 */
static inline or_vec3 conjmat_vec_mul_arg(const or_geom *g, const void *u, int link, int idx_mat,
                                          const void *in, int idx_vect, const double *phases)
{
    or_vec3 o;
    double arg = phases[(size_t)link * g->sizeh + idx_mat];
    or_dc phase = cos(arg) + I * sin(arg);
    or_dc vec0 = vplane(g, in, 0)[idx_vect] * conj(phase);
    or_dc vec1 = vplane(g, in, 1)[idx_vect] * conj(phase);
    or_dc vec2 = vplane(g, in, 2)[idx_vect] * conj(phase);
    or_dc mat00 = uplane(g, u, link, 0, 0)[idx_mat];
    or_dc mat01 = uplane(g, u, link, 0, 1)[idx_mat];
    or_dc mat02 = uplane(g, u, link, 0, 2)[idx_mat];
    or_dc mat10 = uplane(g, u, link, 1, 0)[idx_mat];
    or_dc mat11 = uplane(g, u, link, 1, 1)[idx_mat];
    or_dc mat12 = uplane(g, u, link, 1, 2)[idx_mat];
    or_dc cmat20 = (mat01 * mat12) - (mat02 * mat11);
    or_dc cmat21 = (mat02 * mat10) - (mat00 * mat12);
    or_dc cmat22 = (mat00 * mat11) - (mat01 * mat10);
    o.c0 = (conj(mat00) * vec0) + (conj(mat10) * vec1) + (cmat20 * vec2);
    o.c1 = (conj(mat01) * vec0) + (conj(mat11) * vec1) + (cmat21 * vec2);
    o.c2 = (conj(mat02) * vec0) + (conj(mat12) * vec1) + (cmat22 * vec2);
    return o;
}

/* acc_Deo_unsafe (fermion_matrix.c:47-99) when eo == 0, acc_Doe_unsafe (:101-157) when eo == 1 */
static void dslash(const or_geom *g, const void *u, void *out, const void *in,
                   const double *phases, int eo)
{
    const int nd0 = g->nd[0], nd1 = g->nd[1], nd2 = g->nd[2], nd3 = g->nd[3];
    const int nd0h = nd0 >> 1;
    or_dc *o0 = vplane(g, out, 0), *o1 = vplane(g, out, 1), *o2 = vplane(g, out, 2);
    /* OpenMP only in the CPU-baseline build (oracle/Makefile `bench`); ignored otherwise */
#pragma omp parallel for collapse(2) schedule(static)
    for (int d3 = 0; d3 < nd3; d3++)
        for (int d2 = 0; d2 < nd2; d2++)
            for (int d1 = 0; d1 < nd1; d1++)
                for (int hd0 = 0; hd0 < nd0h; hd0++) {
                    int d1m = d1 - 1; d1m = d1m + (((d1m >> 31) & 0x1) * nd1);
                    int d2m = d2 - 1; d2m = d2m + (((d2m >> 31) & 0x1) * nd2);
                    int d3m = d3 - 1; d3m = d3m + (((d3m >> 31) & 0x1) * nd3);
                    int d1p = d1 + 1; d1p *= (((d1p - nd1) >> 31) & 0x1);
                    int d2p = d2 + 1; d2p *= (((d2p - nd2) >> 31) & 0x1);
                    int d3p = d3 + 1; d3p *= (((d3p - nd3) >> 31) & 0x1);
                    int d0 = 2 * hd0 + ((d1 + d2 + d3 + eo) & 0x1);
                    int d0m = d0 - 1; d0m = d0m + (((d0m >> 31) & 0x1) * nd0);
                    int d0p = d0 + 1; d0p *= (((d0p - nd0) >> 31) & 0x1);
                    or_vec3 aux = {0, 0, 0}, t;
                    int im[4] = {or_snum(g, d0m, d1, d2, d3), or_snum(g, d0, d1m, d2, d3),
                                 or_snum(g, d0, d1, d2m, d3), or_snum(g, d0, d1, d2, d3m)};
                    for (int mu = 0; mu < 4; mu++) {
                        /* Deo: backward links are the odd arrays 2mu+1; Doe: the even arrays 2mu */
                        int link = 2 * mu + (eo == 0 ? 1 : 0);
                        t = conjmat_vec_mul_arg(g, u, link, im[mu], in, im[mu], phases);
                        aux.c0 -= t.c0; aux.c1 -= t.c1; aux.c2 -= t.c2;
                    }
                    int idxh = or_snum(g, d0, d1, d2, d3);
                    int ip[4] = {or_snum(g, d0p, d1, d2, d3), or_snum(g, d0, d1p, d2, d3),
                                 or_snum(g, d0, d1, d2p, d3), or_snum(g, d0, d1, d2, d3p)};
                    for (int mu = 0; mu < 4; mu++) {
                        int link = 2 * mu + (eo == 0 ? 0 : 1);
                        t = mat_vec_mul_arg(g, u, link, idxh, in, ip[mu], phases);
                        aux.c0 += t.c0; aux.c1 += t.c1; aux.c2 += t.c2;
                    }
                    o0[idxh] = (aux.c0) * 0.5;
                    o1[idxh] = (aux.c1) * 0.5;
                    o2[idxh] = (aux.c2) * 0.5;
                }
}

void or_acc_Deo(const or_geom *g, const void *u, void *out, const void *in, const double *phases)
{
    dslash(g, u, out, in, phases, 0);
}
void or_acc_Doe(const or_geom *g, const void *u, void *out, const void *in, const double *phases)
{
    dslash(g, u, out, in, phases, 1);
}

/* FP32 twin: the sp_ build of the same loop (float data, cosf/sinf, float complex) */
typedef struct { or_fc c0, c1, c2; } or_vec3f;

static inline or_vec3f mat_vec_mul_arg_f(const or_geom *g, const void *u, int link, int idx_mat,
                                         const void *in, int idx_vect, const float *phases)
{
    or_vec3f o;
    float arg = phases[(size_t)link * g->sizeh + idx_mat];
    or_fc phase = cosf(arg) + I * sinf(arg);
    or_fc vec0 = vplane_f(g, in, 0)[idx_vect] * phase;
    or_fc vec1 = vplane_f(g, in, 1)[idx_vect] * phase;
    or_fc vec2 = vplane_f(g, in, 2)[idx_vect] * phase;
    or_fc mat00 = uplane_f(g, u, link, 0, 0)[idx_mat];
    or_fc mat01 = uplane_f(g, u, link, 0, 1)[idx_mat];
    or_fc mat02 = uplane_f(g, u, link, 0, 2)[idx_mat];
    or_fc mat10 = uplane_f(g, u, link, 1, 0)[idx_mat];
    or_fc mat11 = uplane_f(g, u, link, 1, 1)[idx_mat];
    or_fc mat12 = uplane_f(g, u, link, 1, 2)[idx_mat];
    or_fc mat20 = conjf((mat01 * mat12) - (mat02 * mat11));
    or_fc mat21 = conjf((mat02 * mat10) - (mat00 * mat12));
    or_fc mat22 = conjf((mat00 * mat11) - (mat01 * mat10));
    o.c0 = (mat00 * vec0) + (mat01 * vec1) + (mat02 * vec2);
    o.c1 = (mat10 * vec0) + (mat11 * vec1) + (mat12 * vec2);
    o.c2 = (mat20 * vec0) + (mat21 * vec1) + (mat22 * vec2);
    return o;
}

static inline or_vec3f conjmat_vec_mul_arg_f(const or_geom *g, const void *u, int link, int idx_mat,
                                             const void *in, int idx_vect, const float *phases)
{
    or_vec3f o;
    float arg = phases[(size_t)link * g->sizeh + idx_mat];
    or_fc phase = cosf(arg) + I * sinf(arg);
    or_fc vec0 = vplane_f(g, in, 0)[idx_vect] * conjf(phase);
    or_fc vec1 = vplane_f(g, in, 1)[idx_vect] * conjf(phase);
    or_fc vec2 = vplane_f(g, in, 2)[idx_vect] * conjf(phase);
    or_fc mat00 = uplane_f(g, u, link, 0, 0)[idx_mat];
    or_fc mat01 = uplane_f(g, u, link, 0, 1)[idx_mat];
    or_fc mat02 = uplane_f(g, u, link, 0, 2)[idx_mat];
    or_fc mat10 = uplane_f(g, u, link, 1, 0)[idx_mat];
    or_fc mat11 = uplane_f(g, u, link, 1, 1)[idx_mat];
    or_fc mat12 = uplane_f(g, u, link, 1, 2)[idx_mat];
    or_fc cmat20 = (mat01 * mat12) - (mat02 * mat11);
    or_fc cmat21 = (mat02 * mat10) - (mat00 * mat12);
    or_fc cmat22 = (mat00 * mat11) - (mat01 * mat10);
    o.c0 = (conjf(mat00) * vec0) + (conjf(mat10) * vec1) + (cmat20 * vec2);
    o.c1 = (conjf(mat01) * vec0) + (conjf(mat11) * vec1) + (cmat21 * vec2);
    o.c2 = (conjf(mat02) * vec0) + (conjf(mat12) * vec1) + (cmat22 * vec2);
    return o;
}

static void dslash_f(const or_geom *g, const void *u, void *out, const void *in,
                     const float *phases, int eo)
{
    const int nd0 = g->nd[0], nd1 = g->nd[1], nd2 = g->nd[2], nd3 = g->nd[3];
    const int nd0h = nd0 >> 1;
    or_fc *o0 = vplane_f(g, out, 0), *o1 = vplane_f(g, out, 1), *o2 = vplane_f(g, out, 2);
    #pragma omp parallel for collapse(2) schedule(static)
    for (int d3 = 0; d3 < nd3; d3++)
        for (int d2 = 0; d2 < nd2; d2++)
            for (int d1 = 0; d1 < nd1; d1++)
                for (int hd0 = 0; hd0 < nd0h; hd0++) {
                    int d1m = d1 - 1; d1m = d1m + (((d1m >> 31) & 0x1) * nd1);
                    int d2m = d2 - 1; d2m = d2m + (((d2m >> 31) & 0x1) * nd2);
                    int d3m = d3 - 1; d3m = d3m + (((d3m >> 31) & 0x1) * nd3);
                    int d1p = d1 + 1; d1p *= (((d1p - nd1) >> 31) & 0x1);
                    int d2p = d2 + 1; d2p *= (((d2p - nd2) >> 31) & 0x1);
                    int d3p = d3 + 1; d3p *= (((d3p - nd3) >> 31) & 0x1);
                    int d0 = 2 * hd0 + ((d1 + d2 + d3 + eo) & 0x1);
                    int d0m = d0 - 1; d0m = d0m + (((d0m >> 31) & 0x1) * nd0);
                    int d0p = d0 + 1; d0p *= (((d0p - nd0) >> 31) & 0x1);
                    or_vec3f aux = {0, 0, 0}, t;
                    int im[4] = {or_snum(g, d0m, d1, d2, d3), or_snum(g, d0, d1m, d2, d3),
                                 or_snum(g, d0, d1, d2m, d3), or_snum(g, d0, d1, d2, d3m)};
                    for (int mu = 0; mu < 4; mu++) {
                        int link = 2 * mu + (eo == 0 ? 1 : 0);
                        t = conjmat_vec_mul_arg_f(g, u, link, im[mu], in, im[mu], phases);
                        aux.c0 -= t.c0; aux.c1 -= t.c1; aux.c2 -= t.c2;
                    }
                    int idxh = or_snum(g, d0, d1, d2, d3);
                    int ip[4] = {or_snum(g, d0p, d1, d2, d3), or_snum(g, d0, d1p, d2, d3),
                                 or_snum(g, d0, d1, d2p, d3), or_snum(g, d0, d1, d2, d3p)};
                    for (int mu = 0; mu < 4; mu++) {
                        int link = 2 * mu + (eo == 0 ? 0 : 1);
                        t = mat_vec_mul_arg_f(g, u, link, idxh, in, ip[mu], phases);
                        aux.c0 += t.c0; aux.c1 += t.c1; aux.c2 += t.c2;
                    }
                    o0[idxh] = (aux.c0) * 0.5f;
                    o1[idxh] = (aux.c1) * 0.5f;
                    o2[idxh] = (aux.c2) * 0.5f;
                }
}

void or_acc_Deo_f(const or_geom *g, const void *u, void *out, const void *in, const float *ph)
{
    dslash_f(g, u, out, in, ph, 0);
}
void or_acc_Doe_f(const or_geom *g, const void *u, void *out, const void *in, const float *ph)
{
    dslash_f(g, u, out, in, ph, 1);
}

/* ------------------------------------------------------------------- BLAS */

double or_l2norm2(const or_geom *g, const void *v)
{
    double resR = 0.0;
#pragma omp parallel for reduction(+ : resR) schedule(static)
    for (long t = 0; t < g->sizeh; t++) {
        /* l2norm2_loc, fermionic_utilities.h:32-40 */
        or_dc a = vplane(g, v, 0)[t], b = vplane(g, v, 1)[t], c = vplane(g, v, 2)[t];
        double sum = creal(a) * creal(a) + cimag(a) * cimag(a);
        sum += creal(b) * creal(b) + cimag(b) * cimag(b);
        sum += creal(c) * creal(c) + cimag(c) * cimag(c);
        resR += sum;
    }
    return resR;
}

double or_real_scal_prod(const or_geom *g, const void *x, const void *y)
{
    double resR = 0.0;
#pragma omp parallel for reduction(+ : resR) schedule(static)
    for (long t = 0; t < g->sizeh; t++) {
        /* scal_prod_loc_1double, fermionic_utilities.h:20-28 */
        double sum = 0;
        for (int c = 0; c < 3; c++) {
            or_dc a = vplane(g, x, c)[t], b = vplane(g, y, c)[t];
            if (c == 0) sum = creal(a) * creal(b) + cimag(a) * cimag(b);
            else sum += creal(a) * creal(b) + cimag(a) * cimag(b);
        }
        resR += sum;
    }
    return resR;
}

or_dc or_scal_prod(const or_geom *g, const void *x, const void *y)
{
    double resR = 0.0, resI = 0.0;
    for (long t = 0; t < g->sizeh; t++) {
        or_dc color_sum = conj(vplane(g, x, 0)[t]) * vplane(g, y, 0)[t];
        color_sum += conj(vplane(g, x, 1)[t]) * vplane(g, y, 1)[t];
        color_sum += conj(vplane(g, x, 2)[t]) * vplane(g, y, 2)[t];
        resR += creal(color_sum);
        resI += cimag(color_sum);
    }
    return resR + resI * I;
}

/* the sp_ reductions keep a double accumulator (fermionic_utilities.c is not
 * in the transformer's all-types list, double_to_single_transformer.py:105-145) */
double or_l2norm2_f(const or_geom *g, const void *v)
{
    double resR = 0.0;
    for (long t = 0; t < g->sizeh; t++) {
        double sum = 0;
        for (int c = 0; c < 3; c++) {
            or_fc a = vplane_f(g, v, c)[t];
            sum += (double)crealf(a) * crealf(a) + (double)cimagf(a) * cimagf(a);
        }
        resR += sum;
    }
    return resR;
}

double or_real_scal_prod_f(const or_geom *g, const void *x, const void *y)
{
    double resR = 0.0;
    for (long t = 0; t < g->sizeh; t++) {
        double sum = 0;
        for (int c = 0; c < 3; c++) {
            or_fc a = vplane_f(g, x, c)[t], b = vplane_f(g, y, c)[t];
            sum += (double)crealf(a) * crealf(b) + (double)cimagf(a) * cimagf(b);
        }
        resR += sum;
    }
    return resR;
}

/* out = in1*factor + in2 (combine_in1xfactor_plus_in2, fermionic_utilities.c:180-192) */
static void axpy_to(const or_geom *g, const void *in1, double factor, const void *in2, void *out)
{
    for (int c = 0; c < 3; c++) {
        or_dc *a = vplane(g, in1, c), *b = vplane(g, in2, c), *o = vplane(g, out, c);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < g->sizeh; i++) o[i] = (a[i] * factor) + b[i];
    }
}
/* sp_ BLAS: float data, double factor (fermionic_utilities.c keeps `double`
 * under the transformer), so the arithmetic is promoted and rounded on store */
static void axpy_to_f(const or_geom *g, const void *in1, double factor, const void *in2, void *out)
{
    for (int c = 0; c < 3; c++) {
        or_fc *a = vplane_f(g, in1, c), *b = vplane_f(g, in2, c), *o = vplane_f(g, out, c);
        for (long i = 0; i < g->sizeh; i++) o[i] = (or_fc)(((or_dc)a[i] * factor) + (or_dc)b[i]);
    }
}
static void sub_to(const or_geom *g, const void *in1, const void *in2, void *out)
{
    for (int c = 0; c < 3; c++) {
        or_dc *a = vplane(g, in1, c), *b = vplane(g, in2, c), *o = vplane(g, out, c);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < g->sizeh; i++) o[i] = a[i] - b[i];
    }
}
static void copy_to(const or_geom *g, const void *in, void *out)
{
    memmove(out, in, or_vec3_bytes(g));
}

/* fermion_matrix_multiplication_shifted (fermion_matrix.c:735-746) with
 * combine_in1xferm_mass_minus_in2 (fermionic_utilities.c:261-272) */
void or_fermion_matrix_multiplication_shifted(const or_geom *g, const void *u, void *out,
        const void *in, void *temp, const double *phases, double mass, double shift)
{
    or_acc_Doe(g, u, temp, in, phases);
    or_acc_Deo(g, u, out, temp, phases);
    double m2 = mass * mass + shift;
    for (int c = 0; c < 3; c++) {
        or_dc *a = vplane(g, in, c), *o = vplane(g, out, c);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < g->sizeh; i++) o[i] = (a[i] * m2) - o[i];
    }
}

void or_fermion_matrix_multiplication_shifted_f(const or_geom *g, const void *u, void *out,
        const void *in, void *temp, const float *phases, double mass, double shift)
{
    or_acc_Doe_f(g, u, temp, in, phases);
    or_acc_Deo_f(g, u, out, temp, phases);
    double m2 = mass * mass + (float)shift;
    for (int c = 0; c < 3; c++) {
        or_fc *a = vplane_f(g, in, c), *o = vplane_f(g, out, c);
        for (long i = 0; i < g->sizeh; i++) o[i] = (or_fc)(((or_dc)a[i] * m2) - (or_dc)o[i]);
    }
}

/* ---------------------------------------------------------------- solvers */

#define OR_SUCCESS 1
#define OR_FAILURE 0

/* ker_invert_openacc, inverter_full.c:19-132 (SAFETY_MARGIN 0.95 at :17) */
int or_ker_invert(const or_geom *g, const void *u, const double *phases, double mass,
                  void *solution, const void *in, double res,
                  void *loc_r, void *loc_h, void *loc_s, void *loc_p, int max_cg, double shift,
                  const or_tricks *tr, int *cg_return)
{
    const double SAFETY_MARGIN = 0.95;
    int cg = 0;
    double delta, alpha, lambda = 0, omega, gammag;
    double source_norm = or_l2norm2(g, in);
    do {
        or_fermion_matrix_multiplication_shifted(g, u, loc_s, solution, loc_h, phases, mass, shift);
        sub_to(g, in, loc_s, loc_r);
        copy_to(g, loc_r, loc_p);
        delta = or_l2norm2(g, loc_r);
        int cg_restarted = 0;
        do {
            cg++;
            cg_restarted++;
            or_fermion_matrix_multiplication_shifted(g, u, loc_s, loc_p, loc_h, phases, mass, shift);
            alpha = or_real_scal_prod(g, loc_p, loc_s);
            omega = delta / alpha;
            axpy_to(g, loc_p, omega, solution, solution);
            axpy_to(g, loc_s, -omega, loc_r, loc_r);
            lambda = or_l2norm2(g, loc_r);
            gammag = lambda / delta;
            delta = lambda;
            axpy_to(g, loc_p, gammag, loc_r, loc_p);
        } while ((sqrt(lambda / source_norm) > res * SAFETY_MARGIN) &&
                 cg_restarted < tr->restarting_every);
    } while ((sqrt(lambda / source_norm) > res) && cg < max_cg);

    or_fermion_matrix_multiplication_shifted(g, u, loc_s, solution, loc_h, phases, mass, shift);
    sub_to(g, in, loc_s, loc_h);
    double current_res = or_l2norm2(g, loc_h) / source_norm;
    *cg_return = cg;
    return (sqrt(current_res) <= res) ? OR_SUCCESS : OR_FAILURE;
}

void or_vec3_d2f(const or_geom *g, const void *d, void *f)
{
    for (int c = 0; c < 3; c++) {
        or_dc *a = vplane(g, d, c); or_fc *b = vplane_f(g, f, c);
        for (long i = 0; i < g->sizeh; i++) b[i] = (float)creal(a[i]) + I * (float)cimag(a[i]);
    }
}
void or_vec3_f2d(const or_geom *g, const void *f, void *d)
{
    for (int c = 0; c < 3; c++) {
        or_fc *a = vplane_f(g, f, c); or_dc *b = vplane(g, d, c);
        for (long i = 0; i < g->sizeh; i++) b[i] = (double)crealf(a[i]) + I * (double)cimagf(a[i]);
    }
}
void or_su3_d2f(const or_geom *g, const void *d, void *f, int nlinks)
{
    for (int l = 0; l < nlinks; l++)
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                or_dc *a = uplane(g, d, l, r, c); or_fc *b = uplane_f(g, f, l, r, c);
                for (long i = 0; i < g->sizeh; i++) b[i] = (float)creal(a[i]) + I * (float)cimag(a[i]);
            }
}

/* inverter_mixed_precision, inverter_mixedp.c:41-181 (SAFETY_MARGIN 0.9 at :39).
 * FP32 iteration on u_f/phases_f, FP64 "magic touch" reliable update. */
int or_inverter_mixed_precision(const or_geom *g, const void *u, const void *u_f,
                  const double *phases, const float *phases_f, double mass,
                  void *solution, const void *in, double res, int max_cg, double shift,
                  const or_tricks *tr, int *cg_return)
{
    const double SAFETY_MARGIN = 0.9;
    size_t vb = or_vec3_bytes(g), vbf = vb / 2;
    void *d_r = malloc(vb), *d_h = malloc(vb), *d_s = malloc(vb);
    void *loc_r = malloc(vbf), *loc_h = malloc(vbf), *loc_s = malloc(vbf), *loc_p = malloc(vbf),
         *out = malloc(vbf);
    int cg;
    double delta, alpha, lambda, omega, gammag, lastMaxResNorm = 0;
    double source_norm = or_l2norm2(g, in);

    or_fermion_matrix_multiplication_shifted(g, u, d_s, solution, d_h, phases, mass, shift);
    sub_to(g, in, d_s, d_r);
    or_vec3_d2f(g, d_r, loc_r);
    memcpy(loc_p, loc_r, vbf);
    delta = or_l2norm2_f(g, loc_r);
    memset(out, 0, vbf);
    cg = 0;
    do {
        cg++;
        or_fermion_matrix_multiplication_shifted_f(g, u_f, loc_s, loc_p, loc_h, phases_f, mass, shift);
        alpha = or_real_scal_prod_f(g, loc_p, loc_s);
        omega = delta / alpha;
        axpy_to_f(g, loc_p, omega, out, out);
        if (lastMaxResNorm < delta) lastMaxResNorm = delta;
        if (delta < tr->mixed_precision_delta * lastMaxResNorm) {
            /* combine_add_in2_into_in1_mixed_precision, inverter_mixedp.c:24-34 */
            for (int c = 0; c < 3; c++) {
                or_dc *a = vplane(g, solution, c); or_fc *b = vplane_f(g, out, c);
                for (long i = 0; i < g->sizeh; i++) a[i] += (double)crealf(b[i]) + I * (double)cimagf(b[i]);
            }
            or_fermion_matrix_multiplication_shifted(g, u, d_s, solution, d_h, phases, mass, shift);
            sub_to(g, in, d_s, d_r);
            or_vec3_d2f(g, d_r, loc_r);
            memset(out, 0, vbf);
            lastMaxResNorm = 0;
        } else {
            axpy_to_f(g, loc_s, -omega, loc_r, loc_r);
        }
        lambda = or_l2norm2_f(g, loc_r);
        gammag = lambda / delta;
        delta = lambda;
        axpy_to_f(g, loc_p, gammag, loc_r, loc_p);
    } while ((sqrt(lambda / source_norm) > res * SAFETY_MARGIN) && cg < max_cg);

    for (int c = 0; c < 3; c++) {
        or_dc *a = vplane(g, solution, c); or_fc *b = vplane_f(g, out, c);
        for (long i = 0; i < g->sizeh; i++) a[i] += (double)crealf(b[i]) + I * (double)cimagf(b[i]);
    }
    or_fermion_matrix_multiplication_shifted(g, u, d_s, solution, d_h, phases, mass, shift);
    sub_to(g, in, d_s, d_h);
    double giustoono = or_l2norm2(g, d_h) / source_norm;
    *cg_return = cg;
    free(d_r); free(d_h); free(d_s); free(loc_r); free(loc_h); free(loc_s); free(loc_p); free(out);
    return (sqrt(giustoono) <= res) ? OR_SUCCESS : OR_FAILURE;
}

/* multishift_invert, inverter_multishift_full.c:23-252 (SAFETY_MARGIN 0.95 at :18) */
int or_multishift_invert(const or_geom *g, const void *u, const double *phases, double mass,
                  int order, const double *RA_b, void *out, const void *in, double residuo,
                  void *loc_r, void *loc_h, void *loc_s, void *loc_p, void *shiftferm, int max_cg,
                  int *cg_return)
{
    const double SAFETY_MARGIN = 0.95;
    const size_t vb = or_vec3_bytes(g);
    double zeta_i[OR_MAX_APPROX_ORDER], zeta_ii[OR_MAX_APPROX_ORDER], zeta_iii[OR_MAX_APPROX_ORDER];
    double omegas[OR_MAX_APPROX_ORDER], gammas[OR_MAX_APPROX_ORDER];
    int flag[OR_MAX_APPROX_ORDER];
    int cg, iter, maxiter = 0;
    double alpha, delta, lambda, omega, omega_save, gammag, fact;
#define OUT(k) ((char *)out + (size_t)(k) * vb)
#define SF(k) ((char *)shiftferm + (size_t)(k) * vb)
    for (iter = 0; iter < order; iter++) {
        flag[iter] = 1;
        memset(OUT(iter), 0, vb);
    }
    copy_to(g, in, loc_r);
    copy_to(g, loc_r, loc_p);
    delta = or_l2norm2(g, loc_r);
    double source_norm = or_l2norm2(g, in);
    omega = 1.0;
    for (iter = 0; iter < order; iter++) {
        copy_to(g, in, SF(iter));
        zeta_i[iter] = 1.0;
        zeta_ii[iter] = 1.0;
        gammas[iter] = 0.0;
    }
    gammag = 0.0;
    cg = 0;
    for (iter = 0; iter < order; iter++)
        if (flag[iter] == 1) maxiter = iter + 1;
    do {
        cg++;
        or_fermion_matrix_multiplication_shifted(g, u, loc_s, loc_p, loc_h, phases, mass, 0.0);
        alpha = or_real_scal_prod(g, loc_p, loc_s);
        omega_save = omega;
        omega = -delta / alpha;
        for (iter = 0; iter < maxiter; iter++) {
            if (flag[iter] == 1) {
                zeta_iii[iter] = (zeta_i[iter] * zeta_ii[iter] * omega_save) /
                    (omega * gammag * (zeta_i[iter] - zeta_ii[iter]) +
                     zeta_i[iter] * omega_save * (1.0 - (RA_b[iter]) * omega));
                omegas[iter] = omega * zeta_iii[iter] / zeta_ii[iter];
            }
        }
        /* multiple_combine_in1_minus_in2x_factor_back_into_in1 (fermionic_utilities.c:315-340) */
        for (iter = 0; iter < maxiter; iter++) {
            if (flag[iter] == 1) {
                double factor = omegas[iter];
                for (int c = 0; c < 3; c++) {
                    or_dc *o = vplane(g, OUT(iter), c), *p = vplane(g, SF(iter), c);
                    for (long i = 0; i < g->sizeh; i++) o[i] -= (factor) * (p[i]);
                }
            }
        }
        /* combine_add_factor_x_in2_to_in1 (:206-217) */
        for (int c = 0; c < 3; c++) {
            or_dc *r = vplane(g, loc_r, c), *s = vplane(g, loc_s, c);
            for (long i = 0; i < g->sizeh; i++) r[i] += (omega) * (s[i]);
        }
        lambda = or_l2norm2(g, loc_r);
        gammag = lambda / delta;
        axpy_to(g, loc_p, gammag, loc_r, loc_p);
        for (iter = 0; iter < order; iter++)
            if (flag[iter] == 1)
                gammas[iter] = gammag * zeta_iii[iter] * omegas[iter] / (zeta_ii[iter] * omega);
        /* multiple1_combine_in1_x_fact1_plus_in2_x_fact2_back_into_in1 (:342-378) */
        for (iter = 0; iter < maxiter; iter++) {
            if (flag[iter] == 1) {
                double fact1 = gammas[iter], fact2 = zeta_iii[iter];
                for (int c = 0; c < 3; c++) {
                    or_dc *p = vplane(g, SF(iter), c), *r = vplane(g, loc_r, c);
                    for (long i = 0; i < g->sizeh; i++) p[i] = fact1 * (p[i]) + fact2 * (r[i]);
                }
            }
        }
        maxiter = 0;
        for (iter = 0; iter < order; iter++) {
            if (flag[iter] == 1) {
                fact = sqrt(delta * zeta_ii[iter] * zeta_ii[iter] / source_norm);
                if (fact < residuo * SAFETY_MARGIN) flag[iter] = 0;
                else maxiter = iter + 1;
                zeta_i[iter] = zeta_ii[iter];
                zeta_ii[iter] = zeta_iii[iter];
            }
        }
        delta = lambda;
    } while (maxiter > 0 && cg < max_cg);

    int check = 1;
    for (iter = 0; iter < order; iter++) {
        copy_to(g, OUT(iter), loc_p);
        or_fermion_matrix_multiplication_shifted(g, u, loc_s, loc_p, loc_h, phases, mass, RA_b[iter]);
        sub_to(g, in, loc_s, loc_h);
        double giustoono = or_l2norm2(g, loc_h) / source_norm;
        check *= (giustoono <= 1) ? 1 : 0;
    }
    *cg_return = cg;
#undef OUT
#undef SF
    return check == 1 ? OR_SUCCESS : OR_FAILURE;
}

void or_recombine(const or_geom *g, const void *in_shifted, const void *in, void *out,
                  int order, double a0, const double *RA_a)
{
    const size_t vb = or_vec3_bytes(g);
    for (long ih = 0; ih < g->sizeh; ih++) {
        for (int c = 0; c < 3; c++) {
            or_dc *o = vplane(g, out, c), *x = vplane(g, in, c);
            o[ih] = x[ih] * a0;
            for (int it = 0; it < order; it++)
                o[ih] += RA_a[it] * vplane(g, (const char *)in_shifted + (size_t)it * vb, c)[ih];
        }
    }
}

/* eo_inversion (Meas/ferm_meas.c:50-71) with the plain FP64 CG of
 * inverter_wrapper (useMixedPrecision = 0): returns the CG status. */
int or_eo_inversion(const or_geom *g, const void *u, const double *phases, double mass, double res, int max_cg,
                    const void *in_e, const void *in_o, void *out_e, void *out_o, void *phi_e, void *phi_o,
                    void *r, void *h, void *s, void *p, const or_tricks *tr, int *cg_return)
{
    const long n3 = 3 * g->sizeh;
    const or_dc *ie = (const or_dc *)in_e, *io = (const or_dc *)in_o;
    or_dc *pe = (or_dc *)phi_e, *po = (or_dc *)phi_o, *oe = (or_dc *)out_e, *oo = (or_dc *)out_o;
    long i;
    or_acc_Deo(g, u, phi_e, in_o, phases);
    for (i = 0; i < n3; i++) pe[i] = mass * ie[i] - pe[i];   /* fermionic_utilities.c:393-395 */
    for (i = 0; i < n3; i++) oe[i] = 0.0;
    int st = or_ker_invert(g, u, phases, mass, out_e, phi_e, res, r, h, s, p, max_cg, 0.0, tr, cg_return);
    or_acc_Doe(g, u, phi_o, out_e, phases);
    const double f = (double)1 / mass;
    for (i = 0; i < n3; i++) oo[i] = f * (io[i] - po[i]);    /* fermionic_utilities.c:413-415 */
    return st;
}

/* ker_find_max_eigenvalue_openacc (OpenAcc/find_min_max.c:21-58): power
 * iterations of M = m^2 - Deo Doe from the vector in loc_p (overwritten),
 * stopping when the norm changes by <= 1e-5 relatively; returns the last
 * norm, the iteration count in *iters.  shift != 0 is the min-eigenvalue
 * variant's operator (find_min_max.c:61-97, fermion_matrix_multiplication_shifted). */
static double power_loop(const or_geom *g, const void *u, const double *phases, double mass, double shift,
                         void *loc_r, void *loc_h, void *loc_p, int *iters)
{
    const long n3 = 3 * g->sizeh;
    or_dc *p = (or_dc *)loc_p, *r = (or_dc *)loc_r;
    double norm = sqrt(or_l2norm2(g, loc_p)), old_norm;
    int loop_count = 0;
    do {
        const double inorm = 1.0 / norm;
        long i;
        for (i = 0; i < n3; i++) p[i] = inorm * p[i];   /* multiply_fermion_x_doublefactor */
        for (i = 0; i < n3; i++) r[i] = p[i];           /* assign_in_to_out */
        old_norm = norm;
        or_fermion_matrix_multiplication_shifted(g, u, loc_p, loc_r, loc_h, phases, mass, shift);
        norm = sqrt(or_l2norm2(g, loc_p));
        old_norm = fabs(old_norm - norm);
        old_norm /= norm;
        loop_count++;
    } while (old_norm > 1.0e-5);
    if (iters) *iters = loop_count;
    return norm;
}

double or_find_max_eigenvalue(const or_geom *g, const void *u, const double *phases, double mass, void *loc_r,
                              void *loc_h, void *loc_p, int *iters)
{
    return power_loop(g, u, phases, mass, 0.0, loc_r, loc_h, loc_p, iters);
}

/* ker_find_min_eigenvalue_openacc (find_min_max.c:61-97): the same loop on
 * (m^2 + (max - m^2) - m^2) - Deo Doe, min = max - norm */
double or_find_min_eigenvalue(const or_geom *g, const void *u, const double *phases, double mass, void *loc_r,
                              void *loc_h, void *loc_p, double max, int *iters)
{
    const double m2 = mass * mass;
    const double delta = max - m2;
    const double norm = power_loop(g, u, phases, mass, delta - m2, loc_r, loc_h, loc_p, iters);
    return max - norm;
}

