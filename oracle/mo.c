/*
 * Oracle: AO -> MO transform of the Cholesky vectors,
 *   Rkab(k, a, b) = sum_{p,q} L^k_{pq} CA(p, a) CB(q, b),
 * restating chol_gammcor_MOTransf_2 / _Step1 / _Step2
 * (Cholesky_Gammcor.f90:143-332).  TEST INFRASTRUCTURE ONLY.
 *
 * The reference processes the a-range in passes of MaxNk columns with
 * a0k = (k-1)*MaxNk + a0 and a1k = min(k*MaxNk, a1) (Cholesky_Gammcor.f90:196-197);
 * for a0 > 1 the columns past min(k*MaxNk, a1) are never visited and keep the
 * zero they were initialised with.  That is reproduced here.
 *
 * Pinned by tests/test_oracle_mo.py: dense NumPy contraction
 * C_A^T L^k C_B with L^k unpacked from Rkpq, for a0 = 1 (full agreement) and
 * the a0 > 1 column mask derived by hand from the pass formulas.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

int64_t or_mo_transform(const or_basis *b, const or_chol_result *res, int64_t nao,
                        const double *CA, int64_t lda, int64_t a0, int64_t a1,
                        const double *CB, int64_t ldb, int64_t b0, int64_t b1,
                        int64_t max_buffer_mb, double *Rkab)
{
    const int64_t nvec = res->nvecs, NA = a1 - a0 + 1, NB = b1 - b0 + 1;
    const int64_t msd = res->max_subset_dim;
    int64_t max_buf = (max_buffer_mb * 1024 * 1024) / 8;
    int64_t maxnk = max_buf / (nvec * nao);
    if (maxnk > NA) maxnk = NA;
    if (!(maxnk > 0)) return 1;  /* "Buffer size too small for Rpq->Rab" */
    int64_t npasses = NA / maxnk + (NA % maxnk > 0);
    /* shell offsets (ShellLoc, 1-based) */
    int64_t *shloc = (int64_t *)malloc(sizeof(int64_t) * (b->nshells + 1));
    int64_t loc = 1;
    for (int64_t s = 0; s < b->nshells; ++s) { shloc[s] = loc; loc += or_nang(b, s); }
    memset(Rkab, 0, sizeof(double) * nvec * NA * NB);
    double *W = (double *)malloc(sizeof(double) * nvec * maxnk * nao);
    for (int64_t k = 1; k <= npasses; ++k) {
        int64_t a0k = (k - 1) * maxnk + a0;
        int64_t a1k = (k * maxnk < a1) ? k * maxnk : a1;
        int64_t nk = a1k - a0k + 1;
        if (nk <= 0) continue;
        memset(W, 0, sizeof(double) * nvec * nk * nao);
        /* Step 1: W(:, a, p) += R(:, pq) CA(q, a); W(:, a, q) += R(:, pq) CA(p, a) */
        #pragma omp parallel
        {
            #pragma omp for schedule(static)
            for (int64_t kk = 0; kk < nvec; ++kk) {
                for (int64_t X = 1; X <= res->nsubsets; ++X) {
                    int64_t s0 = res->subset_bounds[2 * (X - 1)], s1 = res->subset_bounds[2 * (X - 1) + 1];
                    if (s0 <= 0) continue;
                    const double *R = res->rkpq + (size_t)nvec * msd * (X - 1);
                    for (int64_t ab = s0; ab <= s1; ++ab) {
                        int64_t sa = res->shell_pairs[2 * (ab - 1)] - 1, sb = res->shell_pairs[2 * (ab - 1) + 1] - 1;
                        int64_t na = or_nang(b, sa), nb = or_nang(b, sb);
                        int64_t la = shloc[sa], lb = shloc[sb];
                        int64_t pq = res->shell_pair_loc[3 * (ab - 1) + 1] - 1;
                        for (int64_t bb = 0; bb < nb; ++bb)
                            for (int64_t a = (sa == sb ? bb : 0); a < na; ++a) {
                                int64_t p = la + a - 1, q = lb + bb - 1;
                                double r = R[kk + nvec * pq];
                                ++pq;
                                for (int64_t i = 0; i < nk; ++i) {
                                    int64_t col = a0k + i - 1;
                                    W[kk + nvec * (i + nk * p)] += r * CA[q + lda * col];
                                    if (p != q) W[kk + nvec * (i + nk * q)] += r * CA[p + lda * col];
                                }
                            }
                    }
                }
            }
        }
        /* Step 2: Rab(:, a0k:a1k, b) += W(:, :, p) CB(p, b) */
        #pragma omp parallel for schedule(static)
        for (int64_t bcol = b0; bcol <= b1; ++bcol) {
            for (int64_t i = 0; i < nk; ++i) {
                double *out = Rkab + nvec * ((a0k - a0 + i) + NA * (bcol - b0));
                for (int64_t p = 0; p < nao; ++p) {
                    double c = CB[p + ldb * (bcol - 1)];
                    const double *w = W + nvec * (i + nk * p);
                    for (int64_t kk = 0; kk < nvec; ++kk) out[kk] += w[kk] * c;
                }
            }
        }
    }
    free(W);
    free(shloc);
    return 0;
}
