/*
 * CPU oracle for the gammcor-integrals Cholesky/ERI hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference algorithms (Fortran, modrzejewski/gammcor-integrals) used as the
 * checker in tests, in the smoke check and as the CPU baseline leg of the
 * benchmark.  The product path (libgigpu, HIP) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Pinning of each
 * function (what independent check fixes it) is listed in the header comment
 * of the corresponding .c file.
 *
 * Conventions (all follow the reference):
 *   - integers are int64_t, indices passed in from the caller are 1-based
 *     where the reference arrays are 1-based (shell centers, shell params);
 *   - arrays are column-major;
 *   - V(a,b,c,d) = (ab|cd) with a fastest.
 */
#ifndef GAMMCOR_ORACLE_H
#define GAMMCOR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Basis description: the TAOBasis arrays the reference passes to chol2_*
 * (basis_definitions.f90:94-192).  All pointers are borrowed. */
typedef struct {
    int64_t natoms;
    const double *coords;          /* (3, natoms) bohr */
    int64_t nshells;
    const int64_t *shell_centers;  /* (nshells) 1-based atom index */
    const int64_t *shell_params;   /* (nshells) 1-based parameter-set index */
    int64_t nparams;
    const int64_t *momentum;       /* (nparams) */
    const int64_t *nprim;          /* (nparams) */
    int64_t max_nprim;
    const double *cntr;            /* (max_nprim, nparams) primitive-normalized */
    const double *exps;            /* (max_nprim, nparams) */
    int64_t max_nang;
    const double *norm;            /* (max_nang, nparams) NormFactorsSpher or Cart */
    int64_t spher;                 /* 1 = spherical AOs (p stays x,y,z) */
} or_basis;

/* ---- Boys function (boys.f90) ---- */
void or_boys_init(int64_t maxm);
double or_boys_fm_reference(int64_t m, double x);
void or_boys_fm(int64_t m, double x, double *F);           /* F[0..m] */
const double *or_boys_table(int64_t *maxm, int64_t *nint); /* (5, nint, 0:maxm) */

/* ---- Cartesian -> real solid harmonics (AutoSpherTransf.py:62-158) ---- */
/* U[(2l+1) x ncart(l)] row-major: U[m*ncart + c], m = -l..l -> 0..2l */
void or_spher_u(int64_t l, double *U);

/* ---- ERIs (Auto2e) ---- */
int64_t or_nang(const or_basis *b, int64_t shell0);
/* V(a,b,c,d) for 0-based shells; Boys per (e0|f0) block as the reference. */
void or_eri_quartet(const or_basis *b, int64_t sa, int64_t sb, int64_t sc, int64_t sd,
                    double kappa, double *V);
/* Same but with one Boys evaluation (top order la+lb+lc+ld) per primitive
 * quartet; used only to measure the Boys top-order sensitivity. */
/* d(ab|cd)/dA, A the centre of shell sa (only shell sa's functions move):
 * dV[3][na nb nc nd], each block in or_eri_quartet's layout.  Non-zero
 * return when la + 1 + lb or the total order is beyond the oracle's tables. */
int or_eri_deriv_a(const or_basis *b, int64_t sa, int64_t sb, int64_t sc, int64_t sd, double kappa, double *dV);
void or_eri_quartet_single_boys(const or_basis *b, int64_t sa, int64_t sb, int64_t sc,
                                int64_t sd, double kappa, double *V);

/* Dense (pq|rs) over shell pairs (0-based (A,B), A >= B), reference storage
 * order inside each pair; V n x n column-major. */
void or_eri_pair_matrix(const or_basis *b, int64_t npairs, const int64_t *pairs, double kappa, double *V,
                        int64_t n);

/* ---- sorting (common/sort.f90:19-330) ---- */
void or_dsort0(double *dx, int64_t *dy, int64_t n, int64_t kflag);

/* ---- two-step Cholesky (TwoStepCholesky*.f90, Cholesky_Gammcor.f90) ---- */
typedef struct {
    int64_t status;            /* 0 ok, else error code */
    char message[256];
    int64_t nvecs;
    int64_t norb_pairs;        /* BASE orbital pairs */
    int64_t nshell_pairs;      /* BASE shell pairs */
    int64_t nsubsets;
    int64_t max_subset_dim;
    int64_t npivot_shell_pairs;
    int64_t nmacro;            /* macro iterations of step 1 */
    /* arrays, malloc'ed, freed by or_chol_free */
    int64_t *shell_pairs;      /* (2, nshell_pairs) 1-based shells, BASE order */
    int64_t *shell_pair_loc;   /* (3, nshell_pairs) FULL, SUBSET_STORAGE, SUBSET_INDEX */
    int64_t *shell_pair_dim;   /* (nshell_pairs) */
    int64_t *subset_dim;       /* (nsubsets) */
    int64_t *subset_bounds;    /* (2, nsubsets) */
    int64_t *pivots;           /* (norb_pairs) 0/1 */
    int64_t *pivot_order;      /* (nvecs) FULL-storage 1-based location, pivot order */
    int64_t *pivot_shell_pairs;     /* (npivot_shell_pairs) */
    int64_t *pivot_shell_pair_loc;  /* (npivot_shell_pairs) */
    int64_t *pivot_shell_pair_dim;  /* (npivot_shell_pairs) */
    int64_t *pivot_orb_pairs;       /* (nvecs) */
    double *vdiag;             /* (nao(nao+1)/2) */
    double *inv_l;             /* (nvecs, nvecs) */
    double *rkpq;              /* (nvecs, max_subset_dim, nsubsets) or NULL */
    double max_residual;       /* Dmax after step 1 */
    double t_step1, t_step2, t_vectors;
    double t_2e, t_rtr, t_rk, t_compress, t_vpqrs, t_chol, t_inv;
    int64_t nquartets;         /* shell quartets evaluated */
} or_chol_result;

/* flags */
#define OR_CHOL_DECODE_FIXED 1   /* decode pivots with the Step-1 triangle order */
#define OR_CHOL_SKIP_VECTORS 2   /* stop after step 2 (no Rkpq) */
#define OR_CHOL_RAW_TIES 4       /* compare sort keys and the batch argmax on raw
                                    doubles (the reference's comparison) instead
                                    of the quantized keys described in chol.c */

void or_chol_build(const or_basis *b, double tau, double kappa, int64_t max_block_dim,
                   int64_t vec_block, int64_t flags, or_chol_result *res);
void or_chol_free(or_chol_result *res);
/* decode of a lower-triangle compound index (TwoStepCholesky_Step2.f90:422-466) */
void or_pq2p_ge_q(int64_t pq, int64_t n, int64_t *p, int64_t *q);

/* ---- MO transform (Cholesky_Gammcor.f90:143-332) ---- */
/* Rkab(nvecs, na*nb) = sum_pq R(k,pq) CA(p,a) CB(q,b); CA/CB are (nao, ncols)
 * column-major already in Auto2e AO order; a0..a1 / b0..b1 are 1-based. */
int64_t or_mo_transform(const or_basis *b, const or_chol_result *res, int64_t nao,
                        const double *CA, int64_t lda, int64_t a0, int64_t a1,
                        const double *CB, int64_t ldb, int64_t b0, int64_t b1,
                        int64_t max_buffer_mb, double *Rkab);

/* ---- One-electron integrals (onee.c) ---- */
/* S, T, V_ne in the AO basis (nao x nao, leading dim ldh, any may be NULL);
 * V from the atom charges (natoms, dummy atoms 0) and npc point charges. */
void or_one_electron(const or_basis *b, const double *atom_charges, int64_t npc, const double *pc_xyz,
                     const double *pc_q, double *S, double *T, double *V, int64_t ldh);

/* ---- Coulomb energy of an AO density (coulomb.c) ---- */
/* E_J = 1/2 sum D_pq (pq|rs) D_rs; D column-major (nao x nao, leading dim ldd). */
double or_coulomb_energy(const or_basis *b, const double *D, int64_t ldd, double kappa);
/* dE_J/dR (3, natoms) from derivative ERIs, for the atoms listed (all when
 * natoms_sel < 0); non-zero return if a class is beyond the oracle's tables */
int or_coulomb_grad(const or_basis *b, const double *D, int64_t ldd, double kappa, int64_t natoms_sel,
                    const int64_t *atoms, double *G);

/* ---- dense linear algebra used by the oracle (LAPACK/BLAS semantics of
 *      dpotrf "L", dtrtri "L","N" and dgemm, real_linalg.f90:39-72,1079-1361) ---- */
int64_t or_dpotrf_lower(int64_t n, double *A, int64_t lda);
int64_t or_dtrtri_lower(int64_t n, double *A, int64_t lda);
void or_dgemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
              const double *A, int64_t lda, const double *B, int64_t ldb, double beta,
              double *C, int64_t ldc);

#ifdef __cplusplus
}
#endif
#endif
