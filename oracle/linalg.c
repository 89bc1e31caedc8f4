/*
 * Oracle: dense linear algebra with the LAPACK/BLAS semantics the reference
 * obtains from MKL (real_linalg.f90:39-72 dpotrf "L" + zeroed upper triangle,
 * dtrtri "L","N"; real_linalg.f90:1079-1111,1331-1361 dgemm).
 * TEST INFRASTRUCTURE ONLY.  Plain OpenMP C, no external library.
 *
 * Pinned by tests/test_oracle_linalg.py against numpy/scipy (LAPACK) on
 * random SPD matrices: L L^T = A, L^-1 L = I, and GEMM for all four
 * transpose combinations.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* right-looking column Cholesky, lower; returns 0 or the LAPACK info */
int64_t or_dpotrf_lower(int64_t n, double *A, int64_t lda)
{
    for (int64_t j = 0; j < n; ++j) {
        double d = A[j + lda * j];
        if (!(d > 0.0)) return j + 1;
        d = sqrt(d);
        A[j + lda * j] = d;
        for (int64_t i = j + 1; i < n; ++i) A[i + lda * j] /= d;
        #pragma omp parallel for schedule(static)
        for (int64_t c = j + 1; c < n; ++c) {
            double f = A[c + lda * j];
            if (f == 0.0) continue;
            double *colc = A + lda * c;
            const double *colj = A + lda * j;
            for (int64_t i = c; i < n; ++i) colc[i] -= colj[i] * f;
        }
    }
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < j; ++i) A[i + lda * j] = 0.0;
    return 0;
}

/* in-place inverse of a lower triangular matrix (dtrti2 "L","N" order:
 * columns from last to first, using the already inverted trailing block) */
int64_t or_dtrtri_lower(int64_t n, double *A, int64_t lda)
{
    for (int64_t j = 0; j < n; ++j)
        if (A[j + lda * j] == 0.0) return j + 1;
    double *tmp = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int64_t j = n - 1; j >= 0; --j) {
        A[j + lda * j] = 1.0 / A[j + lda * j];
        double ajj = -A[j + lda * j];
        int64_t m = n - j - 1;
        if (m > 0) {
            /* x = X(j+1:n, j+1:n) * A(j+1:n, j), X lower triangular */
            for (int64_t i = 0; i < m; ++i) tmp[i] = A[(j + 1 + i) + lda * j];
            #pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < m; ++i) {
                double s = 0.0;
                for (int64_t q = 0; q <= i; ++q) s += A[(j + 1 + i) + lda * (j + 1 + q)] * tmp[q];
                A[(j + 1 + i) + lda * j] = s * ajj;
            }
        }
    }
    free(tmp);
    return 0;
}

/* C = alpha op(A) op(B) + beta C, column-major */
void or_dgemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
              const double *A, int64_t lda, const double *B, int64_t ldb, double beta,
              double *C, int64_t ldc)
{
    if (m <= 0 || n <= 0) return;
    int tA = (ta == 'T' || ta == 't' || ta == 'C' || ta == 'c');
    int tB = (tb == 'T' || tb == 't' || tb == 'C' || tb == 'c');
    /* At[p + k*i] = op(A)(i,p); Bc[p + k*j] = op(B)(p,j) */
    double *At = (double *)malloc(sizeof(double) * (size_t)k * m + 8);
    double *Bc = (double *)malloc(sizeof(double) * (size_t)k * n + 8);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i)
        for (int64_t p = 0; p < k; ++p)
            At[p + k * i] = tA ? A[p + lda * i] : A[i + lda * p];
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = 0; p < k; ++p)
            Bc[p + k * j] = tB ? B[j + ldb * p] : B[p + ldb * j];
    const int64_t IB = 64;
    #pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i0 = 0; i0 < m; i0 += IB) {
            const double *b = Bc + k * j;
            int64_t i1 = i0 + IB < m ? i0 + IB : m;
            for (int64_t i = i0; i < i1; ++i) {
                const double *a = At + k * i;
                double s = 0.0;
                for (int64_t p = 0; p < k; ++p) s += a[p] * b[p];
                double *c = C + i + ldc * j;
                *c = (beta == 0.0 ? 0.0 : beta * *c) + alpha * s;
            }
        }
    free(At);
    free(Bc);
}
