/*
 * Oracle: Singleton's quicksort (CACM Algorithm 347, 1969) in the variant the
 * reference uses, DSORT0 (SLATEC DSORT), common/sort.f90:19-330.
 * TEST INFRASTRUCTURE ONLY.
 *
 * The pivot order of the two-step Cholesky depends on the exact permutation
 * this unstable sort produces for tied keys (shell pairs with bit-identical
 * diagonal maxima), so the steps below follow the published algorithm one to
 * one: median-of-three on a moving fraction R of the segment, the stack of
 * 21 segments, and straight insertion for segments below the first one
 * (which relies on the element left of the segment acting as a sentinel).
 * kflag = +-2 carries dy along; +-1 sorts dx only; negative = descending.
 *
 * Pinned by tests/test_oracle_sort.py: sortedness on random keys, the
 * permutation being consistent with dy, and exhaustive comparison against an
 * independent line-by-line transcription (in Python) of the Fortran
 * statement-label control flow on inputs with many ties.
 */
#include <stdint.h>
#include <stddef.h>
#include "oracle.h"

void or_dsort0(double *dx, int64_t *dy, int64_t n, int64_t kflag)
{
    if (n <= 1) return;
    const int carry = (kflag == 2 || kflag == -2) && dy != NULL;
    if (kflag <= -1)
        for (int64_t i = 0; i < n; ++i) dx[i] = -dx[i];
    double *X = dx - 1;      /* 1-based views */
    int64_t *Y = carry ? dy - 1 : NULL;
    int64_t il[22], iu[22];
    int64_t m = 1, i = 1, j = n, k, l, ij;
    double r = 0.375, t, tt;
    int64_t ty = 0, tty;

new_fraction:
    if (i == j) goto pop_segment;
    if (r <= 0.5898437) r = r + 3.90625e-2;
    else r = r - 0.21875;
partition:
    k = i;
    ij = i + (int64_t)((double)(j - i) * r);
    t = X[ij];
    if (carry) ty = Y[ij];
    if (X[i] > t) {
        X[ij] = X[i]; X[i] = t; t = X[ij];
        if (carry) { Y[ij] = Y[i]; Y[i] = ty; ty = Y[ij]; }
    }
    l = j;
    if (X[j] < t) {
        X[ij] = X[j]; X[j] = t; t = X[ij];
        if (carry) { Y[ij] = Y[j]; Y[j] = ty; ty = Y[ij]; }
        if (X[i] > t) {
            X[ij] = X[i]; X[i] = t; t = X[ij];
            if (carry) { Y[ij] = Y[i]; Y[i] = ty; ty = Y[ij]; }
        }
    }
    for (;;) {
        do { l = l - 1; } while (X[l] > t);
        do { k = k + 1; } while (X[k] < t);
        if (k > l) break;
        tt = X[l]; X[l] = X[k]; X[k] = tt;
        if (carry) { tty = Y[l]; Y[l] = Y[k]; Y[k] = tty; }
    }
    if (l - i > j - k) {
        il[m] = i; iu[m] = l; i = k; m = m + 1;
    } else {
        il[m] = k; iu[m] = j; j = l; m = m + 1;
    }
    goto check_segment;
pop_segment:
    m = m - 1;
    if (m == 0) goto finish;
    i = il[m];
    j = iu[m];
check_segment:
    if (j - i >= 1) goto partition;
    if (i == 1) goto new_fraction;
    i = i - 1;
    for (;;) {
        i = i + 1;
        if (i == j) goto pop_segment;
        t = X[i + 1];
        if (carry) ty = Y[i + 1];
        if (X[i] <= t) continue;
        k = i;
        do {
            X[k + 1] = X[k];
            if (carry) Y[k + 1] = Y[k];
            k = k - 1;
        } while (t < X[k]);
        X[k + 1] = t;
        if (carry) Y[k + 1] = ty;
    }
finish:
    if (kflag <= -1)
        for (int64_t q = 0; q < n; ++q) dx[q] = -dx[q];
}
