// Rational approximations of x^(num/den) on the host: file names, .REMEZ
// reading and writing, rescaling to the spectral range of M^dag M, evaluation
// (RationalApprox/rationalapprox.c).  The .REMEZ text format and the rescaling
// arithmetic are the reference's, so a host that reads its approximation
// files through this library gets bit-identical RationalApprox structs and
// rescaled coefficients.  Host code only; no device work.
#include "ostp_runtime.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

using namespace ostp;

namespace {

// "approx_<num>_over_<den>" + middle + "_mloglm_<-log10 lambda_min>.REMEZ"
char* approx_name(const char* middle, int exponent_num, int exponent_den, double lambda_min) {
    char buf[128];
    snprintf(buf, sizeof buf, "approx_%d_over_%d%s_mloglm_%1.1f.REMEZ", exponent_num, exponent_den, middle,
             -log(lambda_min) / log(10.0));
    // the reference returns a malloc'ed buffer of 50 characters that the caller frees
    const size_t n = strlen(buf) + 1;
    char* out = (char*)malloc(n < 50 ? 50 : n);
    if (out) memcpy(out, buf, n);
    return out;
}

// one header line "<prefix><value>", matched like fscanf with the same format
bool scan(FILE* f, const char* fmt, int expected, const char* file, const char* what, ...) {
    va_list ap;
    va_start(ap, what);
    const int got = vfscanf(f, fmt, ap);
    va_end(ap);
    if (got == expected) return true;
    fail("%s: not read expected number of entries (%s): %d read vs %d should_read", file, what, got, expected);
    return false;
}

}  // namespace

extern "C" {

// rationalapprox.c:48-72 (by tolerance) and :16-45 (by order)
char* rational_approx_filename(double error, int exponent_num, int exponent_den, double lambda_min) {
    char middle[32];
    snprintf(middle, sizeof middle, "_mlogerr_%1.1lf", -log(error) / log(10.0));
    return approx_name(middle, exponent_num, exponent_den, lambda_min);
}

char* rational_approx_filename_old(int approx_order, int exponent_num, int exponent_den, double lambda_min) {
    char middle[32];
    snprintf(middle, sizeof middle, "_order_%d", approx_order);
    return approx_name(middle, exponent_num, exponent_den, lambda_min);
}

// rationalapprox.c:83-117.  Returns 1 if the file is missing; a malformed file
// is an error (the reference exits).
int rationalapprox_read_custom_nomefile(RationalApprox* ra, char* nomefile) {
    FILE* in = fopen(nomefile, "rt");
    printf("%s\n", nomefile);
    if (!in) {
        printf("Rational Approximation File %s not found!\n", nomefile);
        return 1;
    }
    bool ok = scan(in, "\nApproximation to f(x) = (x)^(%i/%i)\n", 2, nomefile, "exponent", &ra->exponent_num,
                   &ra->exponent_den) &&
              scan(in, "Order: %i\n", 1, nomefile, "order", &ra->approx_order) &&
              scan(in, "Lambda Min: %lf\n", 1, nomefile, "lambda min", &ra->lambda_min) &&
              scan(in, "Lambda Max: %lf\n", 1, nomefile, "lambda max", &ra->lambda_max) &&
              scan(in, "GMP Remez Precision: %i\n", 1, nomefile, "precision", &ra->gmp_remez_precision) &&
              scan(in, "Error: %lf\n", 1, nomefile, "error", &ra->error) &&
              scan(in, "RA_a0 = %lf\n", 1, nomefile, "a0", &ra->RA_a0);
    if (ok && (ra->approx_order < 0 || ra->approx_order > MAX_APPROX_ORDER))
        ok = fail("%s: order %d exceeds %d", nomefile, ra->approx_order, MAX_APPROX_ORDER);
    for (int k = 0; ok && k < ra->approx_order; k++) {
        int i0 = 0, i1 = 0;
        ok = scan(in, "RA_a[%i] = %lf, RA_b[%i] = %lf\n", 4, nomefile, "pole", &i0, &ra->RA_a[k], &i1, &ra->RA_b[k]);
    }
    fclose(in);
    if (!ok) return 2;
    if (verbosity_lv > 4) {
        printf("RA_a0 = %18.16e\n", ra->RA_a0);
        for (int k = 0; k < ra->approx_order; k++)
            printf("RA_a[%d] = %18.16e, RA_b[%d] = %18.16e\n", k, ra->RA_a[k], k, ra->RA_b[k]);
    }
    return 0;
}

// rationalapprox.c:74-81: the file named after error, exponent and lambda_min
int rationalapprox_read(RationalApprox* ra) {
    char* name = rational_approx_filename(ra->error, ra->exponent_num, ra->exponent_den, ra->lambda_min);
    if (!name) return 1;
    const int e = rationalapprox_read_custom_nomefile(ra, name);
    free(name);
    return e;
}

// rationalapprox.c:119-143
void rationalapprox_save(const char* nomefile, RationalApprox* ra) {
    FILE* out = fopen(nomefile, "w");
    printf("%s\n", nomefile);
    if (!out) {
        fail("Could not open file %s", nomefile);
        return;
    }
    fprintf(out, "\nApproximation to f(x) = (x)^(%i/%i)\n", ra->exponent_num, ra->exponent_den);
    fprintf(out, "Order: %i\n", ra->approx_order);
    fprintf(out, "Lambda Min: %18.16e\n", ra->lambda_min);
    fprintf(out, "Lambda Max: %18.16e\n", ra->lambda_max);
    fprintf(out, "GMP Remez Precision: %i\n", ra->gmp_remez_precision);
    fprintf(out, "Error: %18.16e\n", ra->error);
    fprintf(out, "RA_a0 = %18.16e\n", ra->RA_a0);
    for (int k = 0; k < ra->approx_order; k++)
        fprintf(out, "RA_a[%d] = %18.16e, RA_b[%d] = %18.16e\n", k, ra->RA_a[k], k, ra->RA_b[k]);
    fclose(out);
}

// rationalapprox.c:145-194: an approximation on [lambda_min, 1] mapped to
// [0.95 min, 1.05 max] of the measured spectrum; the range check is the
// reference's (fatal if the mother approximation does not reach min, a
// warning if it reaches far below)
void rescale_rational_approximation(RationalApprox* in, RationalApprox* out, double* minmax) {
    const double power = (double)in->exponent_num / in->exponent_den;
    out->exponent_num = in->exponent_num;
    out->exponent_den = in->exponent_den;
    out->approx_order = in->approx_order;
    out->gmp_remez_precision = in->gmp_remez_precision;
    out->error = in->error;
    double mn = minmax[0], mx = minmax[1];
    mn *= 0.95;
    mx *= 1.05;
    (void)mn;
    const double eps = pow(mx, power);
    out->RA_a0 = in->RA_a0 * eps;
    for (int k = 0; k < in->approx_order; k++) {
        out->RA_a[k] = in->RA_a[k] * mx * eps;
        out->RA_b[k] = in->RA_b[k] * mx;
    }
    out->lambda_min = in->lambda_min * mx;
    out->lambda_max = mx;
    if (out->lambda_min > minmax[0]) {
        printf("ERROR: mother rational approx does not cover the range!\n");
        printf("out->lambda_max: %e \n", out->lambda_max);
        printf("out->lambda_min: %e , minmax[0]: %e\n", out->lambda_min, minmax[0]);
        printf("Consider modifying your input file, setting\n");
        printf("ExpMaxEigenvalue         %f     # OR LARGER!!\n", mx * 1.2);
        fail("rescale_rational_approximation: approximation range [%e, %e] does not cover %e", out->lambda_min,
             out->lambda_max, minmax[0]);
    } else if (out->lambda_min < 0.3 * minmax[0]) {
        printf("Warning, the range of your rational approximation is really large\n");
        printf("out->lambda_min: %e , minmax[0]: %e\n", out->lambda_min, minmax[0]);
        printf("You may consider reducing 'ExpMaxEigenvalue' in the input file to %f.\n", mx * 1.2);
    }
}

// rationalapprox.c:197-220: map [lambda_min, lambda_max] to [., 1]
void renormalize_rational_approximation(RationalApprox* in, RationalApprox* out) {
    const double power = (double)in->exponent_num / in->exponent_den;
    out->exponent_num = in->exponent_num;
    out->exponent_den = in->exponent_den;
    out->approx_order = in->approx_order;
    out->gmp_remez_precision = in->gmp_remez_precision;
    out->error = in->error;
    const double ratio = 1 / in->lambda_max;
    const double eps = pow(ratio, power);
    out->RA_a0 = in->RA_a0 * eps;
    for (int k = 0; k < in->approx_order; k++) {
        out->RA_a[k] = in->RA_a[k] * ratio * eps;
        out->RA_b[k] = in->RA_b[k] * ratio;
    }
    out->lambda_min = in->lambda_min * ratio;
    out->lambda_max = 1;
}

// rationalapprox.c:223-234: a0 + sum_k a_k / (x + b_k), summed in order
double rational_approx_evaluate(RationalApprox* in, double x) {
    double r = in->RA_a0;
    for (int k = 0; k < in->approx_order; k++) r += in->RA_a[k] / (x + in->RA_b[k]);
    return r;
}

}  // extern "C"
