/*
 * Oracle: Boys function F_m(x) = Int_0^1 t^(2m) exp(-x t^2) dt, restated from
 * the reference's tabulated evaluator.  TEST INFRASTRUCTURE ONLY.
 *
 *   fm_reference  boys.f90:94-163 (+ Shavitt expansion boys.f90:166-199)
 *   table         boys_init boys.f90:57-86, BOYS_ACCURACY=3 -> X_asym=25,
 *                 500 intervals, 5-point Chebyshev (boys.f90:20-51);
 *                 cheb_nodes / cheb_coeffs chebinterp.f90:9-56
 *   fm            boys.f90:224-281 with fm_interp boys.f90:202-221 and
 *                 cheb_approx_5 chebinterp.f90:122-149
 *
 * Pinned by: tests/test_oracle_boys.py -- mpmath quadrature of the integral
 * definition (relative error of or_boys_fm_reference < 4e-15, of the table
 * < 4.3e-12 as boys.f90:30 states), the closed form F_0 = sqrt(pi/x)erf(sqrt x)/2,
 * the x>=25 branch equal to the truncated asymptotic expression, and the
 * downward-recursion identity between orders.
 */
#include <math.h>
#include <stdlib.h>
#include "oracle.h"

#define OR_PI 3.141592653589793
#define X_ASYM 25.0
#define NINT 500
#define NPT 5

static double *g_table = NULL; /* (NPT, NINT, 0:maxm) */
static int64_t g_maxm = -1;

static double shavitt(int64_t m, double twox, double expmx)
{
    /* boys.f90:166-199: six terms of sum (2x)^i / ((2m+1)(2m+3)...(2m+2i+1)) */
    double sum = 0.0, xi = 1.0;
    int64_t a = 2 * m + 1, d = a;
    for (int i = 0; i < 6; ++i) {
        sum = sum + xi / (double)d;
        xi = xi * twox;
        a = a + 2;
        d = d * a;
    }
    return expmx * sum;
}

double or_boys_fm_reference(int64_t m, double x)
{
    /* boys.f90:94-163 */
    double f;
    if (x < 35.0) {
        int64_t dm;
        if (x < 1.0) dm = 15;
        else if (x < 5.0) dm = 30;
        else if (x < 10.0) dm = 50;
        else if (x < 15.0) dm = 60;
        else if (x < 20.0) dm = 70;
        else if (x < 25.0) dm = 80;
        else if (x < 30.0) dm = 90;
        else dm = 100;
        double expmx = exp(-x), twox = 2.0 * x;
        f = shavitt(m + dm, twox, expmx);
        int64_t a = 2 * (m + dm - 1) + 1;
        for (int64_t l = 1; l <= dm; ++l) {
            f = 1.0 / (double)a * (twox * f + expmx);
            a -= 2;
        }
    } else {
        f = 0.5 * sqrt(OR_PI / x);
        double expmx = exp(-x), beta = 1.0 / (2.0 * x);
        int64_t a = 1;
        for (int64_t l = 2; l <= m + 1; ++l) {
            f = beta * ((double)a * f - expmx);
            a += 2;
        }
    }
    return f;
}

void or_boys_init(int64_t maxm)
{
    /* boys_init boys.f90:57-86 */
    if (g_table && g_maxm >= maxm) return;
    free(g_table);
    g_table = (double *)malloc(sizeof(double) * NPT * NINT * (maxm + 1));
    g_maxm = maxm;
    const double dx = X_ASYM / (double)NINT;
    for (int64_t k = 0; k <= maxm; ++k) {
        for (int64_t i = 1; i <= NINT; ++i) {
            double a = (double)(i - 1) * dx, b = (double)i * dx;
            double xs[NPT], fs[NPT];
            /* cheb_nodes chebinterp.f90:9-30 */
            double w0 = OR_PI / (double)NPT, w1 = (b - a) / 2.0, w2 = (b + a) / 2.0;
            for (int j = 1; j <= NPT; ++j) {
                double t = w0 * ((double)j - 0.5);
                xs[j - 1] = w1 * cos(t) + w2;
            }
            for (int j = 0; j < NPT; ++j) fs[j] = or_boys_fm_reference(k, xs[j]);
            /* cheb_coeffs chebinterp.f90:33-56 */
            double *c = g_table + NPT * ((i - 1) + NINT * k);
            for (int j = 1; j <= NPT; ++j) {
                double s = 0.0, wj = (double)(j - 1);
                for (int kk = 1; kk <= NPT; ++kk) {
                    double t = w0 * wj * ((double)kk - 0.5);
                    s = s + fs[kk - 1] * cos(t);
                }
                c[j - 1] = 2.0 / (double)NPT * s;
            }
        }
    }
}

const double *or_boys_table(int64_t *maxm, int64_t *nint)
{
    *maxm = g_maxm;
    *nint = NINT;
    return g_table;
}

static double fm_interp(double x, int64_t m)
{
    /* fm_interp boys.f90:202-221 + cheb_approx_5 chebinterp.f90:122-149 */
    const double dx = X_ASYM / (double)NINT;
    int64_t k = (int64_t)floor(x / dx);
    double ab_sum = (double)(2 * k + 1) * dx;
    const double *c = g_table + NPT * (k + NINT * m);
    double y = (2.0 * x - ab_sum) / dx;
    double twoy = 2.0 * y;
    double d5 = c[4];
    double d4 = twoy * d5 + c[3];
    double d3 = twoy * d4 - d5 + c[2];
    double d2 = twoy * d3 - d4 + c[1];
    return y * d2 - d3 + 0.5 * c[0];
}

void or_boys_fm(int64_t m, double x, double *F)
{
    /* fm boys.f90:224-281 (f0..f6 boys.f90:284-515 are the same recursion
     * unrolled, with the division by (2l-1) written as a literal) */
    if (x < X_ASYM) {
        double f = fm_interp(x, m);
        F[m] = f;
        double expmx = exp(-x), twox = 2.0 * x;
        double a = (double)(2 * (m - 1) + 1);
        for (int64_t l = m; l >= 1; --l) {
            f = (twox * f + expmx) / a;
            F[l - 1] = f;
            a = a - 2.0;
        }
    } else {
        double f = 0.5 * sqrt(OR_PI / x);
        F[0] = f;
        double expmx = exp(-x), a = 1.0, beta = 1.0 / (2.0 * x);
        for (int64_t l = 2; l <= m + 1; ++l) {
            f = beta * (a * f - expmx);
            a = a + 2.0;
            F[l - 1] = f;
        }
    }
}
