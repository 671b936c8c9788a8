/*
 * CPU oracle, input generators and rational approximations -- TEST
 * INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Generators: the reference draws from a sequential dSFMT stream
 * (Rand/random.c:20-54), which cannot be reproduced on a GPU without running
 * serially.  The product library instead uses a counter-based Philox4x32-10
 * stream keyed by (seed) and addressed by (index, component, stream, tag); the
 * field *rules* are the reference ones (random_assignement.c:29-250).  This
 * file restates the library's generator so tests can check it bit-for-bit on
 * the integer stream and to rounding on the derived doubles.
 *
 * Philox4x32-10 is Salmon et al., "Parallel random numbers: as easy as 1, 2,
 * 3" (SC11); constants from that paper.
 */
#include "oracle.h"
#include "su3.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_TWOPI (2 * 3.14159265358979323846)

enum { TAG_GAUSS = 1, TAG_Z2 = 2, TAG_COLD = 3, TAG_EXPH = 4 };

void or_philox4x32_10(const unsigned int ctr_in[4], const unsigned int key_in[2], unsigned int out[4])
{
    const unsigned int M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    unsigned int c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    unsigned int k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += W0; k1 += W1; }
        unsigned long long p0 = (unsigned long long)M0 * c0;
        unsigned long long p1 = (unsigned long long)M1 * c2;
        unsigned int hi0 = (unsigned int)(p0 >> 32), lo0 = (unsigned int)p0;
        unsigned int hi1 = (unsigned int)(p1 >> 32), lo1 = (unsigned int)p1;
        unsigned int n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void draw(unsigned long long seed, unsigned int a, unsigned int b, unsigned int c,
                 unsigned int d, unsigned int out[4])
{
    unsigned int ctr[4] = {a, b, c, d};
    unsigned int key[2] = {(unsigned int)(seed & 0xffffffffu), (unsigned int)(seed >> 32)};
    or_philox4x32_10(ctr, key, out);
}

/* 53-bit uniform in [0,1) from two words */
static double u53(unsigned int x0, unsigned int x1)
{
    return ((double)(x0 >> 5) * 67108864.0 + (double)(x1 >> 6)) * (1.0 / 9007199254740992.0);
}

/* complex Gaussian with the reference's normalisation E|z|^2 = 1
 * (two_double_gauss, random_assignement.c:44-54: radius = sqrt(-log u)) */
static or_dc cgauss(const unsigned int w[4])
{
    double u1 = u53(w[0], w[1]), u2 = u53(w[2], w[3]);
    double phi = OR_TWOPI * u2;
    double radius = sqrt(-1.0 * log(1.0 - u1));
    return radius * cos(phi) + I * (radius * sin(phi));
}

static inline or_dc *vpl(const or_geom *g, void *v, int c)
{
    return (or_dc *)((char *)v + (size_t)c * 16 * g->sizeh);
}

/* generate_vec3_soa_gauss rule (random_assignement.c:67-92) */
void or_gen_vec3_gauss(const or_geom *g, void *v, unsigned long long seed, unsigned int stream)
{
    unsigned int w[4];
    for (long i = 0; i < g->sizeh; i++)
        for (int c = 0; c < 3; c++) {
            draw(seed, (unsigned int)i, (unsigned int)c, stream, TAG_GAUSS, w);
            vpl(g, v, c)[i] = cgauss(w);
        }
}

/* generate_vec3_soa_z2noise rule (random_assignement.c:215-250) */
void or_gen_vec3_z2(const or_geom *g, void *v, unsigned long long seed, unsigned int stream)
{
    unsigned int w[4];
    for (long i = 0; i < g->sizeh; i++)
        for (int c = 0; c < 3; c++) {
            draw(seed, (unsigned int)i, (unsigned int)c, stream, TAG_Z2, w);
            double p = u53(w[0], w[1]);
            vpl(g, v, c)[i] = (p < 0.5) ? (1.0 + 0.0 * I) : (-1.0 + 0.0 * I);
        }
}

static inline or_dc *upl(const or_geom *g, void *u, int link, int r, int c)
{
    return (or_dc *)((char *)u + (size_t)link * or_su3_bytes(g) + (size_t)(3 * r + c) * 16 * g->sizeh);
}

static double norm3(or_dc a, or_dc b, or_dc c)
{
    double norm = creal(a) * creal(a) + cimag(a) * cimag(a);
    norm += creal(b) * creal(b) + cimag(b) * cimag(b);
    norm += creal(c) * creal(c) + cimag(c) * cimag(c);
    return norm;
}

static void store_su3(const or_geom *g, void *u, int link, long i, const or_su3 *m)
{
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) upl(g, u, link, r, c)[i] = m->comp[r][c];
    if (g->par_temp) { /* K = 1 (no defect), HPT_utilities.c:131 */
        double *K = (double *)((char *)u + (size_t)link * or_su3_bytes(g) + (size_t)144 * g->sizeh);
        K[i] = 1.0;
    }
}

/* generate_random_su3 rule (random_assignement.c:106-142): entries
 * delta_rc + eps*((1-2u) + i(1-2u')) for rows 0,1, Gram-Schmidt, row 2 rebuilt.
 * Entry e = 3r+c uses draw(i, 8*link+e, 0, TAG_COLD). */
void or_gen_conf_cold(const or_geom *g, void *u, double f, unsigned long long seed)
{
    unsigned int w[4];
    for (int link = 0; link < 8; link++)
        for (long i = 0; i < g->sizeh; i++) {
            or_su3 m;
            for (int e = 0; e < 6; e++) {
                draw(seed, (unsigned int)i, (unsigned int)(8 * link + e), 0, TAG_COLD, w);
                double a = 1.0 - 2.0 * u53(w[0], w[1]);
                double b = 1.0 - 2.0 * u53(w[2], w[3]);
                or_dc z = (a + I * b) * f;
                m.comp[e / 3][e % 3] = (e % 4 == 0) ? 1 + z : z; /* diagonal e = 0, 4 */
            }
            double norm = norm3(m.comp[0][0], m.comp[0][1], m.comp[0][2]);
            norm = 1.0 / sqrt(norm);
            m.comp[0][0] *= norm; m.comp[0][1] *= norm; m.comp[0][2] *= norm;
            or_dc prod = conj(m.comp[0][0]) * m.comp[1][0] + conj(m.comp[0][1]) * m.comp[1][1] +
                         conj(m.comp[0][2]) * m.comp[1][2];
            m.comp[1][0] -= prod * m.comp[0][0];
            m.comp[1][1] -= prod * m.comp[0][1];
            m.comp[1][2] -= prod * m.comp[0][2];
            norm = norm3(m.comp[1][0], m.comp[1][1], m.comp[1][2]);
            norm = 1.0 / sqrt(norm);
            m.comp[1][0] *= norm; m.comp[1][1] *= norm; m.comp[1][2] *= norm;
            or_rebuild3row(&m);
            store_su3(g, u, link, i, &m);
        }
}

/* U = exp(i*beta*H), H traceless Hermitian with standard-normal entries:
 * 5 draws (i, 8*link+k, 0, TAG_EXPH), k=0..4, give 10 normals n0..n9 by
 * Box-Muller (sqrt(-2 log(1-u1)) cos/sin 2 pi u2).  Diagonal h_kk = n_k -
 * (n0+n1+n2)/3; H01 = n3 + i n4, H02 = n5 + i n6, H12 = n7 + i n8.
 * exp(i beta H) = exp(-QA) with QA = -i beta H, evaluated by the
 * Cayley-Hamilton form of cayley_hamilton.h:52-183. */
void or_gen_conf_expH(const or_geom *g, void *u, double beta, unsigned long long seed)
{
    unsigned int w[4];
    for (int link = 0; link < 8; link++)
        for (long i = 0; i < g->sizeh; i++) {
            double n[10];
            for (int k = 0; k < 5; k++) {
                draw(seed, (unsigned int)i, (unsigned int)(8 * link + k), 0, TAG_EXPH, w);
                double u1 = u53(w[0], w[1]), u2 = u53(w[2], w[3]);
                double rad = sqrt(-2.0 * log(1.0 - u1));
                double phi = OR_TWOPI * u2;
                n[2 * k] = rad * cos(phi);
                n[2 * k + 1] = rad * sin(phi);
            }
            double tr3 = (n[0] + n[1] + n[2]) / 3.0;
            double h00 = n[0] - tr3, h11 = n[1] - tr3;
            or_tamat QA;
            /* QA = -i beta H: off-diagonal -i*beta*H_rc; diagonal i*(-beta*h_kk) */
            QA.c01 = -I * beta * (n[3] + I * n[4]);
            QA.c02 = -I * beta * (n[5] + I * n[6]);
            QA.c12 = -I * beta * (n[7] + I * n[8]);
            QA.ic00 = -beta * h00;
            QA.ic11 = -beta * h11;
            or_su3 m;
            or_ch_exp_minus_QA(&QA, &m);
            or_rebuild3row(&m);
            store_su3(g, u, link, i, &m);
        }
}

/* ------------------------------------------------- rational approximations */

/* rationalapprox_read_custom_nomefile, rationalapprox.c:83-117; returns 0 on
 * success, 1 if the file is missing, 2 on a malformed file (the reference
 * exits the process there, CHECKREAD at :9-13). */
int or_ratapprox_read(or_ratapprox *ra, const char *nomefile)
{
    FILE *input = fopen(nomefile, "rt");
    if (!input) return 1;
#define RD(expr, n) do { if ((expr) != (n)) { fclose(input); return 2; } } while (0)
    RD(fscanf(input, "\nApproximation to f(x) = (x)^(%i/%i)\n", &ra->exponent_num, &ra->exponent_den), 2);
    RD(fscanf(input, "Order: %i\n", &ra->approx_order), 1);
    if (ra->approx_order < 0 || ra->approx_order > OR_MAX_APPROX_ORDER) { fclose(input); return 2; }
    RD(fscanf(input, "Lambda Min: %lf\n", &ra->lambda_min), 1);
    RD(fscanf(input, "Lambda Max: %lf\n", &ra->lambda_max), 1);
    RD(fscanf(input, "GMP Remez Precision: %i\n", &ra->gmp_remez_precision), 1);
    RD(fscanf(input, "Error: %lf\n", &ra->error), 1);
    RD(fscanf(input, "RA_a0 = %lf\n", &ra->RA_a0), 1);
    for (int i = 0; i < ra->approx_order; i++) {
        int i1, i2;
        RD(fscanf(input, "RA_a[%i] = %lf, RA_b[%i] = %lf\n", &i1, &ra->RA_a[i], &i2, &ra->RA_b[i]), 4);
    }
#undef RD
    fclose(input);
    return 0;
}

/* rational_approx_evaluate, rationalapprox.c:225-236 */
double or_ratapprox_evaluate(const or_ratapprox *in, double x)
{
    double res = in->RA_a0;
    for (int order = 0; order < in->approx_order; ++order) res += in->RA_a[order] / (x + in->RA_b[order]);
    return res;
}

/* rescale_rational_approximation, rationalapprox.c:145-194 (without the exit) */
void or_ratapprox_rescale(const or_ratapprox *in, or_ratapprox *out, const double minmax[2])
{
    double power = (double)in->exponent_num / in->exponent_den;
    out->exponent_num = in->exponent_num;
    out->exponent_den = in->exponent_den;
    out->approx_order = in->approx_order;
    out->gmp_remez_precision = in->gmp_remez_precision;
    out->error = in->error;
    double min = minmax[0], max = minmax[1];
    min *= 0.95;
    (void)min;
    max *= 1.05;
    double epsilon = pow(max, power);
    out->RA_a0 = in->RA_a0 * epsilon;
    for (int order = 0; order < in->approx_order; order++) {
        out->RA_a[order] = in->RA_a[order] * max * epsilon;
        out->RA_b[order] = in->RA_b[order] * max;
    }
    out->lambda_min = in->lambda_min * max;
    out->lambda_max = max;
}
