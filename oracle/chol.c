/*
 * Oracle: two-step Cholesky decomposition of the Coulomb matrix V(pq,rs),
 * restating the reference (Folkestad, Kjonstad, Koch, JCP 150, 194112 (2019)).
 * TEST INFRASTRUCTURE ONLY.
 *
 *   chol_gammcor_CoulombMatrix        Cholesky_Gammcor.f90:380-441
 *   chol2_Step1_FullInterface         TwoStepCholesky.f90:74-167
 *   chol2_Vdiag / write_Vdiag         TwoStepCholesky_Step1.f90:175-279
 *   chol2_DefineBase (incl. the kept  TwoStepCholesky_Step1.f90:282-442
 *     last discarded shell pair)
 *   chol2_DefineSubsets/DivideOrbPairs TwoStepCholesky_Step1.f90:445-530
 *   chol2_DefineCandidates            TwoStepCholesky_Step1.f90:533-575
 *   chol2_DefineBatch                 TwoStepCholesky_Step1.f90:578-625
 *   chol2_Compress                    TwoStepCholesky_Step1.f90:16-97
 *   chol2_SubtractR                   TwoStepCholesky_Step1.f90:628-691
 *   chol2_BatchMax (strict >)         TwoStepCholesky_Step1.f90:694-730
 *   chol2_M + write_V* folding        TwoStepCholesky_Step1.f90:100-172,733-856
 *   chol2_Lj                          TwoStepCholesky_Step1.f90:859-881
 *   chol2_Pivots (macro/micro loop)   TwoStepCholesky_Step1.f90:884-1103
 *   chol2_Step2                       TwoStepCholesky.f90:243-364
 *   chol2_PivotShellPairs + decode    TwoStepCholesky_Step2.f90:422-534
 *   chol2_Vpqrs                       TwoStepCholesky_Step2.f90:16-32,76-239
 *   chol2_Vabrs                       TwoStepCholesky_Step2.f90:35-73,242-419
 *   chol2_FullDimVectors_Batch        TwoStepCholesky.f90:378-416
 *
 * Pinned by:
 *   tests/test_oracle_pivots.py   the pivot sequence, BASE shell pairs and
 *                                 vector count equal a dense NumPy restatement
 *                                 of Step 1 written from the Fortran
 *                                 (tests/dense_pivots.py), on uncontracted,
 *                                 contracted, prescreened and 3240-pair
 *                                 systems, quantized and raw keys;
 *   tests/test_oracle_chol.py     R^T R reproduces the factorized columns of
 *                                 V; V - R^T R within tau (fixed decode);
 *                                 Inv_L is the inverse Cholesky factor of
 *                                 V(P,P); subsets do not change the pivots;
 *                                 tau = 1e-10 converges below tau;
 *   tests/test_oracle_step2_mo.py the pq decode against the N=3 traversal
 *                                 drawn at TwoStepCholesky_Step2.f90:432-436;
 *                                 the MO transform against C^T L^k C.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>
#include "oracle.h"

enum { BASE = 1, CAND = 2, BATCH = 3 };
enum { FULL = 1, SSTORE = 2, SINDEX = 3, COMPR = 4 };

static double wall(void) { return omp_get_wtime(); }

/* Near-tie canonicalization.  Keys that are equal in exact arithmetic (e.g.
 * shell pairs mirrored by a molecular symmetry) come out of floating-point
 * evaluation differing in the last bits, with a sign that depends on the
 * implementation's rounding.  The reference orders such keys by that noise.
 * To make the pivot sequence a function of the mathematics (and identical
 * between this oracle and the GPU library), keys are quantized before they
 * are compared:
 *   - diagonal maxima for the BASE sort: mantissa rounded to KEY_REL_BITS
 *     (26) bits;
 *   - residual diagonals (candidate/batch sorts, batch argmax): rounded to
 *     an absolute quantum delta = 2^(floor(log2 MaxDrs) - KEY_ABS_BITS),
 *     KEY_ABS_BITS = 32, or 2^(floor(log2 tau) - KEY_TAU_BITS), KEY_TAU_BITS
 *     = 8, whichever is smaller, so that the quantum stays far below tau.
 * Only the orderings use the quantized keys: the thresholds (tau for the
 * candidates, Dmin for the batch and the micro iterations, tau for
 * convergence) compare the raw values, as the reference does.
 * The quanta sit above the disagreement of two correct ERI evaluations
 * (different Boys-function evaluation and summation order reach 1e-11
 * relative on small diagonals) and far below tau.
 * Quantized ties fall back to the sort's tie order and the first-in-batch
 * rule.  OR_CHOL_RAW_TIES restores raw comparisons. */
enum { KEY_REL_BITS = 26, KEY_ABS_BITS = 32, KEY_TAU_BITS = 8 };

static double qrel(double x)
{
    if (x == 0.0 || !isfinite(x)) return x;
    int e;
    double m = frexp(x, &e);
    return ldexp(rint(ldexp(m, KEY_REL_BITS + 1)), e - KEY_REL_BITS - 1);
}

static double qabs(double x, double inv_delta) { return rint(x * inv_delta); }

typedef struct {
    const or_basis *b;
    double kappa;
    int64_t nq;  /* quartet counter */
} ctx_t;

/* shell pair dimension: Na*Nb or Na(Na+1)/2 for a same-shell pair */
static int64_t pair_dim(const or_basis *b, int64_t sa, int64_t sb)
{
    int64_t na = or_nang(b, sa), nb = or_nang(b, sb);
    return sa != sb ? na * nb : na * (na + 1) / 2;
}

/* copy the (ab|cd) block V(Na,Nb,Nc,Nd) into M(p0.., q0..) with the
 * same-shell lower-triangle folding of rows (A==B) and columns (C==D)
 * (chol2_write_Vabcd/Vabcc/Vaacd/Vaacc, TwoStepCholesky_Step1.f90:100-172) */
static void write_block(double *M, int64_t ldm, int64_t p0, int64_t q0, const double *V,
                        int64_t na, int64_t nb, int64_t nc, int64_t nd, int same_ab, int same_cd)
{
    int64_t q = 0;
    for (int64_t d = 0; d < nd; ++d)
        for (int64_t c = (same_cd ? d : 0); c < nc; ++c) {
            int64_t p = 0;
            for (int64_t bb = 0; bb < nb; ++bb)
                for (int64_t a = (same_ab ? bb : 0); a < na; ++a) {
                    M[(p0 + p) + ldm * (q0 + q)] = V[a + na * (bb + nb * (c + nc * d))];
                    ++p;
                }
            ++q;
        }
}

static int64_t maxnang_of(const or_basis *b)
{
    int64_t lmax = 0;
    for (int64_t p = 0; p < b->nparams; ++p) if (b->momentum[p] > lmax) lmax = b->momentum[p];
    return b->spher ? 2 * lmax + 1 : (lmax + 1) * (lmax + 2) / 2;
}

static void set_error(or_chol_result *res, int64_t code, const char *msg)
{
    res->status = code;
    snprintf(res->message, sizeof(res->message), "%s", msg);
}

void or_pq2p_ge_q(int64_t pq, int64_t n, int64_t *p, int64_t *q)
{
    /* TwoStepCholesky_Step2.f90:422-466 */
    int64_t q_base = (pq - 1) / (n + 1) + 1;
    int64_t v = pq - (n + 1) * (q_base - 1);
    int64_t interval1 = n - q_base + 1;
    int64_t in2 = v / (interval1 + 1);
    int64_t in1 = in2 ^ 1;
    *p = in1 * (q_base + v - 1) + in2 * (v - interval1 + n - q_base);
    *q = in1 * q_base + in2 * interval1;
}

void or_chol_free(or_chol_result *r)
{
    free(r->shell_pairs); free(r->shell_pair_loc); free(r->shell_pair_dim);
    free(r->subset_dim); free(r->subset_bounds); free(r->pivots); free(r->pivot_order);
    free(r->pivot_shell_pairs); free(r->pivot_shell_pair_loc); free(r->pivot_shell_pair_dim);
    free(r->pivot_orb_pairs); free(r->vdiag); free(r->inv_l); free(r->rkpq);
    memset(r, 0, sizeof(*r));
}

void or_chol_build(const or_basis *b, double tau, double kappa, int64_t max_block_dim,
                   int64_t block_dim, int64_t flags, or_chol_result *res)
{
    memset(res, 0, sizeof(*res));
    ctx_t cx = {b, kappa, 0};
    const int64_t nshells = b->nshells;
    int64_t lmax = 0;
    for (int64_t p = 0; p < b->nparams; ++p) if (b->momentum[p] > lmax) lmax = b->momentum[p];
    or_boys_init(4 * 5);
    const int64_t maxnang = maxnang_of(b);
    int64_t nao = 0;
    for (int64_t s = 0; s < nshells; ++s) nao += or_nang(b, s);
    const int64_t max_sp = nshells * (nshells + 1) / 2;
    const int64_t max_op = nao * (nao + 1) / 2;
    double t0 = wall();

    /* ---------------- chol2_Vdiag ---------------- */
    double *vdiag = (double *)malloc(sizeof(double) * max_op);
    res->vdiag = vdiag;
    int64_t *vd_loc = (int64_t *)malloc(sizeof(int64_t) * (max_sp + 1));
    int64_t *vd_dim = (int64_t *)malloc(sizeof(int64_t) * (max_sp + 1));
    int64_t *vd_pair = (int64_t *)malloc(sizeof(int64_t) * 2 * (max_sp + 1));
    double *vd_max = (double *)malloc(sizeof(double) * (max_sp + 1));
    int64_t *idxmap = (int64_t *)malloc(sizeof(int64_t) * (max_sp + 1));
    {
        int64_t loc = 1, ab = 1;
        for (int64_t sb = 0; sb < nshells; ++sb)
            for (int64_t sa = sb; sa < nshells; ++sa) {
                vd_loc[ab] = loc;
                vd_dim[ab] = pair_dim(b, sa, sb);
                vd_pair[2 * ab] = sa + 1;
                vd_pair[2 * ab + 1] = sb + 1;
                loc += vd_dim[ab];
                ++ab;
            }
        #pragma omp parallel for schedule(dynamic, 4)
        for (int64_t k = 1; k <= max_sp; ++k) {
            int64_t sa = vd_pair[2 * k] - 1, sb = vd_pair[2 * k + 1] - 1;
            int64_t na = or_nang(b, sa), nb = or_nang(b, sb);
            double *V = (double *)malloc(sizeof(double) * na * nb * na * nb);
            or_eri_quartet(b, sa, sb, sa, sb, kappa, V);
            int64_t pq = vd_loc[k] - 1;
            for (int64_t bb = 0; bb < nb; ++bb)
                for (int64_t a = (sa == sb ? bb : 0); a < na; ++a)
                    vdiag[pq++] = V[a + na * (bb + nb * (a + na * bb))];
            free(V);
        }
        cx.nq += max_sp;
    }

    /* ---------------- chol2_DefineBase ---------------- */
    for (int64_t k = 1; k <= max_sp; ++k) {
        double mx = -INFINITY;
        for (int64_t i = 0; i < vd_dim[k]; ++i) if (vdiag[vd_loc[k] - 1 + i] > mx) mx = vdiag[vd_loc[k] - 1 + i];
        vd_max[k] = mx;
        idxmap[k] = k;
    }
    double maxdrs = -INFINITY;
    for (int64_t k = 1; k <= max_sp; ++k) if (vd_max[k] > maxdrs) maxdrs = vd_max[k];
    if (maxdrs <= 0.0) { set_error(res, 10, "Maximum diagonal Coulomb integral is nonpositive"); goto cleanup0; }
    double pthresh = fmin(tau, 1.0e-8);
    pthresh = pthresh * pthresh / maxdrs;
    const int raw_ties = (flags & OR_CHOL_RAW_TIES) != 0;
    int maxdrs_exp;
    frexp(maxdrs, &maxdrs_exp);
    int tau_exp;
    frexp(tau, &tau_exp);
    /* 1 / delta, delta = min(2^(floor(log2 MaxDrs) - KEY_ABS_BITS), 2^(floor(log2 tau) - KEY_TAU_BITS)) */
    const double inv_delta = ldexp(1.0, (KEY_ABS_BITS + 1 - maxdrs_exp) > (KEY_TAU_BITS + 1 - tau_exp)
                                            ? (KEY_ABS_BITS + 1 - maxdrs_exp) : (KEY_TAU_BITS + 1 - tau_exp));
    if (!raw_ties)
        for (int64_t k = 1; k <= max_sp; ++k) vd_max[k] = qrel(vd_max[k]);
    or_dsort0(vd_max + 1, idxmap + 1, max_sp, 2);
    int64_t k0 = 1;
    for (int64_t k = 1; k <= max_sp; ++k) {
        int64_t ab = idxmap[k];
        double dpq = -INFINITY;
        for (int64_t i = 0; i < vd_dim[ab]; ++i) if (vdiag[vd_loc[ab] - 1 + i] > dpq) dpq = vdiag[vd_loc[ab] - 1 + i];
        if (dpq <= pthresh) k0 = k; else break;
    }
    const int64_t nsp = max_sp - k0 + 1;
    int64_t nop = 0;
    for (int64_t k = k0; k <= max_sp; ++k) nop += vd_dim[idxmap[k]];
    int64_t *sp = (int64_t *)malloc(sizeof(int64_t) * 2 * (nsp + 1));
    int64_t *sploc = (int64_t *)calloc(5 * (nsp + 1), sizeof(int64_t));   /* [5*s + mode] */
    int64_t *spdim = (int64_t *)malloc(sizeof(int64_t) * (nsp + 1));
    double *D = (double *)malloc(sizeof(double) * (nop + 1));
    {
        int64_t s = 1, p = 1;
        for (int64_t k = k0; k <= max_sp; ++k) {
            int64_t l = idxmap[k];
            sp[2 * s] = vd_pair[2 * l];
            sp[2 * s + 1] = vd_pair[2 * l + 1];
            spdim[s] = vd_dim[l];
            sploc[5 * s + FULL] = p;
            for (int64_t i = 0; i < vd_dim[l]; ++i) D[p + i] = vdiag[vd_loc[l] - 1 + i];
            ++s;
            p += vd_dim[l];
        }
    }
    res->nshell_pairs = nsp;
    res->norb_pairs = nop;

    /* ---------------- chol2_DefineSubsets / DivideOrbPairs ---------------- */
    int64_t nsub = (int64_t)ceil((double)nop / (double)max_block_dim);
    int64_t *sdim = (int64_t *)calloc(nsub + 1, sizeof(int64_t));
    int64_t *sbnd = (int64_t *)calloc(2 * (nsub + 1), sizeof(int64_t));
    {
        int64_t target = (int64_t)ceil((double)nop / (double)nsub);
        int64_t berr[2] = {0, 0}, bloc[2] = {0, 0};
        for (int64_t k = 1; k <= nsub; ++k) {
            if (k == 1) sbnd[2 * k] = 1;
            else sbnd[2 * k] = (sbnd[2 * (k - 1) + 1] < nsp && sbnd[2 * (k - 1) + 1] > 0) ? sbnd[2 * (k - 1) + 1] + 1 : 0;
            sdim[k] = 0;
            sbnd[2 * k + 1] = sbnd[2 * k];
            if (sbnd[2 * k] > 0) {
                if (k < nsub) {
                    berr[0] = nop; berr[1] = nop;
                    for (int64_t s = sbnd[2 * k]; s <= nsp; ++s) {
                        sploc[5 * s + SSTORE] = sdim[k] + 1;
                        sploc[5 * s + SINDEX] = k;
                        sdim[k] += spdim[s];
                        if (sdim[k] >= target) { bloc[1] = s; berr[1] = sdim[k] - target; break; }
                        else { bloc[0] = s; berr[0] = sdim[k] - target; }
                    }
                    if (llabs(berr[0]) < llabs(berr[1])) { sbnd[2 * k + 1] = bloc[0]; sdim[k] = berr[0] + target; }
                    else { sbnd[2 * k + 1] = bloc[1]; sdim[k] = berr[1] + target; }
                } else {
                    sbnd[2 * k + 1] = nsp;
                    for (int64_t s = sbnd[2 * k]; s <= nsp; ++s) {
                        sploc[5 * s + SSTORE] = sdim[k] + 1;
                        sploc[5 * s + SINDEX] = k;
                        sdim[k] += spdim[s];
                    }
                }
            }
        }
    }
    int64_t max_subset_dim = 0;
    for (int64_t k = 1; k <= nsub; ++k) if (sdim[k] > max_subset_dim) max_subset_dim = sdim[k];
    res->nsubsets = nsub;
    res->max_subset_dim = max_subset_dim;
    const int64_t max_nchol = (int64_t)fmin(ceil((double)nao * 15.0), (double)nop);

    /* ---------------- chol2_Pivots ---------------- */
    int64_t *pivots = (int64_t *)calloc(nop + 1, sizeof(int64_t));
    int64_t *pivot_order = (int64_t *)calloc(max_nchol + 1, sizeof(int64_t));
    int64_t nchol = 0;
    int converged = 0;
    double dmax = 0.0;
    {
        int64_t nsps[4] = {0, nsp, nsp, 0}, nops[4] = {0, nop, nop, 0};
        int64_t *spmap = (int64_t *)calloc(3 * (nsp + 1) + 3, sizeof(int64_t)); /* spmap[k + (nsp+1)*(mode-1)] */
#define SPMAP(k, mode) spmap[(k) + (nsp + 1) * ((mode) - 1)]
        int64_t *colloc = (int64_t *)calloc(nsp + 1, sizeof(int64_t));
        double *spmax = (double *)calloc(nsp + 1, sizeof(double));   /* raw residual maximum per shell pair */
        int64_t max_nblocks = max_nchol / block_dim + (max_nchol % block_dim > 0);
        double **Vb = (double **)calloc(max_nblocks + 1, sizeof(double *));
        for (int64_t k = 1; k <= nsp; ++k) { SPMAP(k, CAND) = k; sploc[5 * k + COMPR] = sploc[5 * k + FULL]; }
        dmax = -INFINITY;
        for (int64_t i = 1; i <= nop; ++i) if (D[i] > dmax) dmax = D[i];
        int64_t cur_block = 0, block_boundary = 0;
        int do_macro = max_nchol > 0;
        double *Dc = D;  /* current compressed D, 1-based */
        int64_t nD = nop; /* its length */
        while (do_macro) {
            res->nmacro++;
            double tc = wall();
            /* chol2_DefineCandidates */
            {
                int64_t n = nsps[CAND];
                double *mv = (double *)malloc(sizeof(double) * (n + 1));
                int64_t *im = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
                for (int64_t k = 1; k <= n; ++k) {
                    int64_t ab = SPMAP(k, CAND), loc = sploc[5 * ab + COMPR];
                    double mx = -INFINITY;
                    for (int64_t i = 0; i < spdim[ab]; ++i) if (Dc[loc + i] > mx) mx = Dc[loc + i];
                    spmax[ab] = mx;
                    mv[k] = raw_ties ? mx : qabs(mx, inv_delta) / inv_delta;
                    im[k] = ab;
                }
                or_dsort0(mv + 1, im + 1, n, -2);
                nsps[CAND] = 0;
                nops[CAND] = 0;
                /* sorted by the (quantized) key, thresholded on the raw maximum:
                 * within one quantum the raw values need not be monotone, so
                 * every pair is tested (raw keys: the reference's loop up to
                 * its first miss) */
                for (int64_t k = 1; k <= n; ++k) {
                    if (spmax[im[k]] > tau) {
                        nsps[CAND]++;
                        SPMAP(nsps[CAND], CAND) = im[k];
                        nops[CAND] += spdim[im[k]];
                    } else if (raw_ties) break;
                }
                free(mv); free(im);
            }
            /* chol2_Compress */
            if (nsps[CAND] > 0) {
                int64_t nc = nsps[CAND];
                int64_t *newloc = (int64_t *)malloc(sizeof(int64_t) * (nc + 1));
                int64_t *oldloc = (int64_t *)malloc(sizeof(int64_t) * (nc + 1));
                int64_t ncomp = 0;
                for (int64_t g = 1; g <= nc; ++g) {
                    int64_t ab = SPMAP(g, CAND);
                    newloc[g] = ncomp + 1;
                    oldloc[g] = sploc[5 * ab + COMPR];
                    ncomp += spdim[ab];
                }
                double *Dn = (double *)malloc(sizeof(double) * (ncomp + 1));
                for (int64_t g = 1; g <= nc; ++g) {
                    int64_t ab = SPMAP(g, CAND);
                    memcpy(Dn + newloc[g], Dc + oldloc[g], sizeof(double) * spdim[ab]);
                }
                if (Dc != D) free(Dc); else free(D);
                D = NULL;
                Dc = Dn;
                nD = ncomp;
                if (nchol > 0) {
                    int64_t nbl = nchol / block_dim + (nchol % block_dim > 0);
                    for (int64_t bl = 1; bl <= nbl; ++bl) {
                        double *T = Vb[bl];
                        double *L = (double *)malloc(sizeof(double) * block_dim * ncomp);
                        #pragma omp parallel for schedule(static)
                        for (int64_t g = 1; g <= nc; ++g) {
                            int64_t ab = SPMAP(g, CAND);
                            memcpy(L + block_dim * (newloc[g] - 1), T + block_dim * (oldloc[g] - 1),
                                   sizeof(double) * block_dim * spdim[ab]);
                        }
                        free(T);
                        Vb[bl] = L;
                    }
                }
                for (int64_t g = 1; g <= nc; ++g) sploc[5 * SPMAP(g, CAND) + COMPR] = newloc[g];
                free(newloc); free(oldloc);
            }
            const int64_t ncand = nops[CAND];
            int64_t max_nbatch = ncand < (maxnang * maxnang > 1000 ? maxnang * maxnang : 1000) ? ncand : (maxnang * maxnang > 1000 ? maxnang * maxnang : 1000);
            res->t_compress += wall() - tc;
            double dmin = fmax(0.01 * dmax, tau);
            double *M = (double *)calloc((size_t)(ncand > 0 ? ncand : 1) * (max_nbatch > 0 ? max_nbatch : 1), sizeof(double));
            double *Lp = (double *)calloc((size_t)(ncand > 0 ? ncand : 1) * (max_nbatch > 0 ? max_nbatch : 1), sizeof(double));
            /* chol2_DefineBatch */
            {
                int64_t n = nsps[CAND];
                double *ds = (double *)malloc(sizeof(double) * (n + 1));
                int64_t *di = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
                for (int64_t k = 1; k <= n; ++k) {
                    int64_t ab = SPMAP(k, CAND), loc = sploc[5 * ab + COMPR];
                    double mx = -INFINITY;
                    for (int64_t i = 0; i < spdim[ab]; ++i) if (Dc[loc + i] > mx) mx = Dc[loc + i];
                    spmax[ab] = mx;
                    ds[k] = raw_ties ? mx : qabs(mx, inv_delta) / inv_delta;
                    di[k] = ab;
                }
                or_dsort0(ds + 1, di + 1, n, -2);
                nsps[BATCH] = 0;
                nops[BATCH] = 0;
                for (int64_t k = 1; k <= n; ++k) {
                    int64_t ab = di[k];
                    if (spmax[ab] > dmin && nops[BATCH] + spdim[ab] <= max_nbatch) {
                        int64_t loc = nops[BATCH] + 1;
                        nsps[BATCH] = k;
                        nops[BATCH] += spdim[ab];
                        SPMAP(k, BATCH) = ab;
                        colloc[k] = loc;
                    } else break;
                }
                free(ds); free(di);
            }
            /* chol2_M */
            double t2 = wall();
            {
                const int64_t nbsp = nsps[BATCH], ncsp = nsps[CAND];
                #pragma omp parallel for collapse(2) schedule(guided)
                for (int64_t L = 1; L <= nbsp; ++L)
                    for (int64_t K = 1; K <= ncsp; ++K) {
                        int64_t ab = SPMAP(K, CAND), cd = SPMAP(L, BATCH);
                        int64_t sa = sp[2 * ab] - 1, sb = sp[2 * ab + 1] - 1;
                        int64_t sc = sp[2 * cd] - 1, sd = sp[2 * cd + 1] - 1;
                        int64_t na = or_nang(b, sa), nb = or_nang(b, sb), nc = or_nang(b, sc), nd = or_nang(b, sd);
                        double *V = (double *)malloc(sizeof(double) * na * nb * nc * nd);
                        or_eri_quartet(b, sa, sb, sc, sd, kappa, V);
                        write_block(M, ncand, sploc[5 * ab + COMPR] - 1, colloc[L] - 1, V, na, nb, nc, nd,
                                    sa == sb, sc == sd);
                        free(V);
                    }
                cx.nq += nbsp * ncsp;
            }
            res->t_2e += wall() - t2;
            /* chol2_SubtractR */
            double tr = wall();
            {
                int64_t nbl = nchol / block_dim + (nchol % block_dim > 0);
                if (nbl > 0 && nops[BATCH] > 0) {
                    double *LB = (double *)malloc(sizeof(double) * block_dim * nops[BATCH]);
                    for (int64_t bl = 1; bl <= nbl; ++bl) {
                        for (int64_t k = 1; k <= nsps[BATCH]; ++k) {
                            int64_t ab = SPMAP(k, BATCH);
                            memcpy(LB + block_dim * (colloc[k] - 1), Vb[bl] + block_dim * (sploc[5 * ab + COMPR] - 1),
                                   sizeof(double) * block_dim * spdim[ab]);
                        }
                        int64_t J = bl < nbl ? block_dim : nchol - (bl - 1) * block_dim;
                        or_dgemm('C', 'N', ncand, nops[BATCH], J, -1.0, Vb[bl], block_dim, LB, block_dim,
                                 1.0, M, ncand);
                    }
                    free(LB);
                }
            }
            res->t_rtr += wall() - tr;
            /* chol2_BatchMax */
            double maxdpq = -1.0;
            int64_t maxloc = 0, maxloc_batch = 0, maxloc_full = 0;
#define BATCHMAX()                                                                 \
            do {                                                                   \
                maxdpq = -1.0;                                                     \
                double bestq = raw_ties ? -1.0 : qabs(-1.0, inv_delta);            \
                for (int64_t k = 1; k <= nsps[BATCH]; ++k) {                       \
                    int64_t ab = SPMAP(k, BATCH), loc = sploc[5 * ab + COMPR];     \
                    for (int64_t p = 1; p <= spdim[ab]; ++p) {                     \
                        double dpq = Dc[loc + p - 1];                              \
                        double dq = raw_ties ? dpq : qabs(dpq, inv_delta);         \
                        if (dq > bestq) {                                          \
                            bestq = dq;                                            \
                            maxdpq = dpq;                                          \
                            maxloc = loc + p - 1;                                  \
                            maxloc_full = sploc[5 * ab + FULL] + p - 1;            \
                            maxloc_batch = colloc[k] + p - 1;                      \
                        }                                                          \
                    }                                                              \
                }                                                                  \
            } while (0)
            BATCHMAX();
            int64_t j = 0;
            int do_micro = nchol < max_nchol && nops[BATCH] > 0;
            double trk = wall();
            while (do_micro) {
                ++j;
                ++nchol;
                if (nchol > block_boundary) {
                    ++cur_block;
                    Vb[cur_block] = (double *)calloc((size_t)block_dim * ncand, sizeof(double));
                    block_boundary = cur_block * block_dim;
                }
                pivots[maxloc_full] = 1;
                pivot_order[nchol] = maxloc_full;
                /* chol2_Lj: L(:,j) = M(:,pj) - L(:,1:j-1) L(pq_j,1:j-1); / sqrt(Dmax) */
                double *Lj = Lp + (size_t)ncand * (j - 1);
                const double *Mj = M + (size_t)ncand * (maxloc_batch - 1);
                const int64_t rowj = maxloc - 1;
                const double sq = sqrt(maxdpq);
                #pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < ncand; ++i) {
                    double s = Mj[i];
                    for (int64_t q = 0; q < j - 1; ++q) s -= Lp[i + (size_t)ncand * q] * Lp[rowj + (size_t)ncand * q];
                    Lj[i] = s / sq;
                }
                double *Vcur = Vb[cur_block];
                int64_t row = nchol - block_dim * (cur_block - 1) - 1;
                for (int64_t i = 0; i < ncand; ++i) {
                    Vcur[row + block_dim * i] = Lj[i];
                    Dc[i + 1] = Dc[i + 1] - Lj[i] * Lj[i];
                }
                BATCHMAX();
                if (maxdpq <= dmin) do_micro = 0;
                if (j == nops[BATCH] || nchol == max_nchol) do_micro = 0;
            }
            res->t_rk += wall() - trk;
            /* raw residual maximum over the rows of D still held (no candidate
             * left: the previous candidates, all <= tau) */
            dmax = -INFINITY;
            for (int64_t i = 1; i <= nD; ++i) if (Dc[i] > dmax) dmax = Dc[i];
            if (nD == 0) dmax = 0.0;
            if (dmax <= tau) { converged = 1; do_macro = 0; }
            if (nchol == max_nchol || nops[BATCH] == 0) do_macro = 0;
            free(M);
            free(Lp);
        }
#undef BATCHMAX
        for (int64_t bl = 1; bl <= max_nblocks; ++bl) free(Vb[bl]);
        free(Vb);
        free(colloc);
        free(spmax);
        free(spmap);
#undef SPMAP
        if (Dc != D) free(Dc);
        else free(D);
    }
    res->max_residual = dmax;
    res->nvecs = nchol;
    res->t_step1 = wall() - t0;
    /* outputs of step 1 */
    res->shell_pairs = (int64_t *)malloc(sizeof(int64_t) * 2 * nsp);
    res->shell_pair_loc = (int64_t *)malloc(sizeof(int64_t) * 3 * nsp);
    res->shell_pair_dim = (int64_t *)malloc(sizeof(int64_t) * nsp);
    for (int64_t s = 1; s <= nsp; ++s) {
        res->shell_pairs[2 * (s - 1)] = sp[2 * s];
        res->shell_pairs[2 * (s - 1) + 1] = sp[2 * s + 1];
        for (int m = 0; m < 3; ++m) res->shell_pair_loc[3 * (s - 1) + m] = sploc[5 * s + 1 + m];
        res->shell_pair_dim[s - 1] = spdim[s];
    }
    res->subset_dim = (int64_t *)malloc(sizeof(int64_t) * nsub);
    res->subset_bounds = (int64_t *)malloc(sizeof(int64_t) * 2 * nsub);
    for (int64_t k = 1; k <= nsub; ++k) {
        res->subset_dim[k - 1] = sdim[k];
        res->subset_bounds[2 * (k - 1)] = sbnd[2 * k];
        res->subset_bounds[2 * (k - 1) + 1] = sbnd[2 * k + 1];
    }
    res->pivots = (int64_t *)malloc(sizeof(int64_t) * nop);
    for (int64_t i = 1; i <= nop; ++i) res->pivots[i - 1] = pivots[i];
    res->pivot_order = (int64_t *)malloc(sizeof(int64_t) * (nchol > 0 ? nchol : 1));
    for (int64_t i = 1; i <= nchol; ++i) res->pivot_order[i - 1] = pivot_order[i];
    free(pivot_order);
    if (nchol == max_nchol && !converged) {
        set_error(res, 20, "Max number of Cholesky vectors reached without convergence");
        free(pivots); free(sp); free(sploc); free(spdim); free(sdim); free(sbnd);
        goto cleanup0;
    }

    /* ---------------- chol2_Step2 ---------------- */
    {
        double ts2 = wall();
        const int64_t nvec = nchol;
        int64_t *psp = (int64_t *)malloc(sizeof(int64_t) * (nsp + 1));
        int64_t *psploc = (int64_t *)malloc(sizeof(int64_t) * (nsp + 1));
        int64_t *pspdim = (int64_t *)malloc(sizeof(int64_t) * (nsp + 1));
        int64_t *porb = (int64_t *)malloc(sizeof(int64_t) * (nvec + 1));
        int64_t npsp = 0, q = 0;
        int64_t nsum = 0;
        for (int64_t i = 1; i <= nop; ++i) nsum += pivots[i];
        if (nsum != nvec) { set_error(res, 30, "Invalid Pivots array"); }
        for (int64_t ab = 1; ab <= nsp; ++ab) {
            int64_t sa = sp[2 * ab], sb = sp[2 * ab + 1];
            int64_t p0 = sploc[5 * ab + FULL], p1 = p0 + spdim[ab] - 1;
            int64_t s = 0;
            for (int64_t p = p0; p <= p1; ++p) s += pivots[p];
            if (s > 0) {
                ++npsp;
                psp[npsp] = ab;
                psploc[npsp] = q + 1;
                pspdim[npsp] = s;
                if (sa != sb) {
                    for (int64_t p = p0; p <= p1; ++p)
                        if (pivots[p] == 1) porb[++q] = p - p0 + 1;
                } else {
                    int64_t na = or_nang(b, sa - 1);
                    for (int64_t p = p0; p <= p1; ++p)
                        if (pivots[p] == 1) {
                            int64_t ab_ = p - p0 + 1, a, bb;
                            if (flags & OR_CHOL_DECODE_FIXED) {
                                /* Step-1 storage order: column-major lower triangle */
                                int64_t cnt = 0; a = bb = 0;
                                for (int64_t y = 1; y <= na && !a; ++y)
                                    for (int64_t x = y; x <= na; ++x)
                                        if (++cnt == ab_) { a = x; bb = y; break; }
                            } else {
                                or_pq2p_ge_q(ab_, na, &a, &bb);
                            }
                            porb[++q] = a + (bb - 1) * na;
                        }
                }
            }
        }
        res->npivot_shell_pairs = npsp;
        /* chol2_Vpqrs */
        double tv = wall();
        double *Vpq = (double *)calloc((size_t)nvec * nvec + 1, sizeof(double));
        const int64_t ntri = npsp * (npsp + 1) / 2;
        #pragma omp parallel for schedule(guided)
        for (int64_t kl = 1; kl <= ntri; ++kl) {
            int64_t K, L;
            or_pq2p_ge_q(kl, npsp, &K, &L);
            int64_t ab = psp[K], cd = psp[L];
            int64_t sa = sp[2 * ab] - 1, sb = sp[2 * ab + 1] - 1, sc = sp[2 * cd] - 1, sd = sp[2 * cd + 1] - 1;
            int64_t na = or_nang(b, sa), nb = or_nang(b, sb), nc = or_nang(b, sc), nd = or_nang(b, sd);
            double *V = (double *)malloc(sizeof(double) * na * nb * nc * nd);
            or_eri_quartet(b, sa, sb, sc, sd, kappa, V);
            int64_t pq0 = psploc[K], rs0 = psploc[L];
            for (int64_t rs = 0; rs < pspdim[L]; ++rs)
                for (int64_t pq = 0; pq < pspdim[K]; ++pq)
                    Vpq[(pq0 - 1 + pq) + nvec * (rs0 - 1 + rs)] =
                        V[(porb[pq0 + pq] - 1) + na * nb * (porb[rs0 + rs] - 1)];
            free(V);
        }
        cx.nq += ntri;
        res->t_vpqrs = wall() - tv;
        double tch = wall();
        int64_t info = or_dpotrf_lower(nvec, Vpq, nvec);
        res->t_chol = wall() - tch;
        if (info != 0) {
            char msg[128];
            snprintf(msg, sizeof(msg), "Cholesky factorization failed with exit code info=%lld", (long long)info);
            set_error(res, 40, msg);
        } else {
            double ti = wall();
            info = or_dtrtri_lower(nvec, Vpq, nvec);
            res->t_inv = wall() - ti;
            if (info != 0) set_error(res, 41, "Lower triangular inverse failed");
        }
        res->inv_l = Vpq;
        res->t_step2 = wall() - ts2;
        /* pivot metadata out */
        res->pivot_shell_pairs = (int64_t *)malloc(sizeof(int64_t) * (npsp > 0 ? npsp : 1));
        res->pivot_shell_pair_loc = (int64_t *)malloc(sizeof(int64_t) * (npsp > 0 ? npsp : 1));
        res->pivot_shell_pair_dim = (int64_t *)malloc(sizeof(int64_t) * (npsp > 0 ? npsp : 1));
        for (int64_t k = 1; k <= npsp; ++k) {
            res->pivot_shell_pairs[k - 1] = psp[k];
            res->pivot_shell_pair_loc[k - 1] = psploc[k];
            res->pivot_shell_pair_dim[k - 1] = pspdim[k];
        }
        res->pivot_orb_pairs = (int64_t *)malloc(sizeof(int64_t) * (nvec > 0 ? nvec : 1));
        for (int64_t i = 1; i <= nvec; ++i) res->pivot_orb_pairs[i - 1] = porb[i];

        /* ---------------- full-dimension vectors ---------------- */
        if (res->status == 0 && !(flags & OR_CHOL_SKIP_VECTORS)) {
            double tfv = wall();
            size_t rk_size = (size_t)nvec * max_subset_dim * nsub;
            res->rkpq = (double *)calloc(rk_size + 1, sizeof(double));
            double *W = (double *)malloc(sizeof(double) * ((size_t)max_subset_dim * nvec + 1));
            for (int64_t X = 1; X <= nsub; ++X) {
                memset(W, 0, sizeof(double) * (size_t)max_subset_dim * nvec);
                int64_t s0 = sbnd[2 * X], s1 = sbnd[2 * X + 1];
                if (s0 > 0) {
                    int64_t nrow = s1 - s0 + 1;
                    #pragma omp parallel for collapse(2) schedule(guided)
                    for (int64_t L = 1; L <= npsp; ++L)
                        for (int64_t r = 0; r < nrow; ++r) {
                            int64_t ab = s0 + r, cd = psp[L];
                            int64_t sa = sp[2 * ab] - 1, sb = sp[2 * ab + 1] - 1, sc = sp[2 * cd] - 1, sd = sp[2 * cd + 1] - 1;
                            int64_t na = or_nang(b, sa), nb = or_nang(b, sb), nc = or_nang(b, sc), nd = or_nang(b, sd);
                            double *V = (double *)malloc(sizeof(double) * na * nb * nc * nd);
                            or_eri_quartet(b, sa, sb, sc, sd, kappa, V);
                            int64_t ab0 = sploc[5 * ab + SSTORE] - 1, rs0 = psploc[L];
                            for (int64_t rs = 0; rs < pspdim[L]; ++rs) {
                                int64_t col = porb[rs0 + rs] - 1;
                                int64_t qq = 0;
                                for (int64_t bb = 0; bb < nb; ++bb)
                                    for (int64_t a = (sa == sb ? bb : 0); a < na; ++a) {
                                        W[(ab0 + qq) + max_subset_dim * (rs0 - 1 + rs)] = V[a + na * bb + na * nb * col];
                                        ++qq;
                                    }
                            }
                            free(V);
                        }
                    cx.nq += nrow * npsp;
                }
                /* Rkpq(:,:,X) = Inv_L * W^T (real_abT) */
                or_dgemm('N', 'T', nvec, max_subset_dim, nvec, 1.0, Vpq, nvec, W, max_subset_dim, 0.0,
                         res->rkpq + (size_t)nvec * max_subset_dim * (X - 1), nvec);
            }
            free(W);
            res->t_vectors = wall() - tfv;
        }
        free(psp); free(psploc); free(pspdim); free(porb);
    }
    free(pivots); free(sp); free(sploc); free(spdim); free(sdim); free(sbnd);
cleanup0:
    res->nquartets = cx.nq;
    free(vd_loc); free(vd_dim); free(vd_pair); free(vd_max); free(idxmap);
}
