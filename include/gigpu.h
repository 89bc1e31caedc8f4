/*
 * gigpu -- MI355X (gfx950) engine for the gammcor-integrals hot path:
 * two-step Cholesky decomposition of the Coulomb matrix (pq|rs) and the
 * AO -> MO transform of its vectors.
 *
 * C ABI only: plain pointers, int64_t sizes, column-major arrays, int status
 * (0 = success; otherwise gi_last_error() describes the failure).  Index
 * arrays carry the 1-based values of the reference's Fortran arrays, so a
 * Fortran caller passes TAOBasis / TChol2Vecs components unchanged through
 * ISO_C_BINDING.  All calls are synchronous with respect to the host on
 * return.  One context per (thread of control, device); a context is not
 * re-entrant, like the reference's module state (Auto2eERI pointer table,
 * Boys INTERPTABLE).
 */
#ifndef GIGPU_H
#define GIGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gi_context gi_context;

/* Accuracy levels of chol_gammcor_Rkpq -> tau
 * (Cholesky_Gammcor.f90:14-17,398-410). */
#define GI_CHOL_ACCURACY_DEFAULT 1   /* tau = 1e-5  */
#define GI_CHOL_ACCURACY_TIGHT 2     /* tau = 1e-6  */
#define GI_CHOL_ACCURACY_LUDICROUS 3 /* tau = 1e-7  */
#define GI_CHOL_ACCURACY_DEBUG 4     /* tau = 1e-10 */

/* gi_chol_build flags */
#define GI_CHOL_DECODE_FIXED 1  /* Step 2 decodes same-shell pivots in the Step-1
                                   storage order instead of chol2_pq2p_ge_q's
                                   (TwoStepCholesky_Step2.f90:520-530) */
#define GI_CHOL_SKIP_VECTORS 2  /* stop after Step 2 (no Rkpq) */
#define GI_CHOL_TIME_ERI 4      /* time ERI launches with HIP events */
#define GI_CHOL_RAW_TIES 8      /* order the sorts and the batch argmax by raw
                                   doubles, as the reference does, instead of by
                                   keys quantized to 2^-26 relative (BASE sort) and
                                   to min(2^-32 of max(Vdiag), 2^-8 of tau)
                                   (residual diagonals); thresholds (tau, Dmin)
                                   always compare raw values */
#define GI_CHOL_SPLIT_BF16 16   /* precision form: the Step-1 trailing update
                                   (chol2_SubtractR) as split-BF16 MFMA products
                                   (four 8-bit slices per element, exact FP32
                                   sums per 256 k, FP64 across chunks; error
                                   ~2^-33 of the operand scales) instead of FP64
                                   dgemm */

/* gi_mo_transform flags */
#define GI_MO_ALL_COLUMNS 1     /* transform every a column even when a0 > 1
                                   (the reference's pass bounds,
                                   Cholesky_Gammcor.f90:196-197, skip some) */
#define GI_MO_SPLIT_BF16 2      /* precision form: the second transform step
                                   (Rab = W CB) as split-BF16 MFMA products (as
                                   GI_CHOL_SPLIT_BF16) instead of FP64 dgemm */

/* Create a context on HIP device `device` (no reference counterpart: the
 * reference initialises module state in chol_gammcor_Rkpq,
 * Cholesky_Gammcor.f90:347-352).  A context keeps its device work arrays
 * (the Cholesky panels, the vector staging, the factorization scratch and
 * the split-product slices) across calls, grown only, so that repeated
 * builds allocate nothing; gi_destroy frees them. */
int gi_create(gi_context **ctx, int64_t device);
int gi_destroy(gi_context *ctx);
/* Message of the last failure on this thread (replaces msg(...,MSG_ERROR) +
 * error stop of the reference).  Status codes, by subsystem:
 *    1  invalid device index             2  HIP / rocBLAS / rocSOLVER call failed
 *    3  shell pair beyond the ERI kernels 4  basis not set
 *    5  block sizes not positive         6  tau not positive
 *    7  null context                     8  invalid argument (shells, pairs, sizes)
 *    9  l beyond 5, or unknown option   10  nonpositive maximum diagonal integral
 *   20  vector limit hit before tau     21  batch too large for the micro iterations
 *   30  invalid pivot array             40  Cholesky factorization of V(P,P) failed
 *   50  no decomposition (build first)  51  empty or invalid Rkab range
 *   52  CA/CB leading dimension < nao   53  buffer too small for Rpq -> Rab
 *   60  no ERI kernel for the orders    61  null density or energy pointer
 *   62  gradient beyond its l limit     63  invalid split-product arguments
 *   64  one-electron integrals, l > 3   65  leading dimension < nao
 *   70  unknown AO ordering             71  AO map not a permutation
 *   72  in-place AO reordering          73  AO reordering beyond the scratch limit
 *   74  scratch array not set up        80  grid orbitals with l > 5
 *   81  invalid grid size / leading dim 82  a grid point where every orbital vanishes
 *   83  THC metric S singular           84  invalid QR threshold or block size
 *   85  grid pruning did not converge   99  C++ exception (message attached)
 * The message text names the failing check; where the reference stops with a
 * message of its own, the text is that message. */
const char *gi_last_error(void);

/* Upload the basis: the TAOBasis components (basis_definitions.f90:94-192)
 * the reference passes to chol2_Step1_FullInterface
 * (TwoStepCholesky.f90:74-105).  coords (3,natoms) bohr; shell_centers,
 * shell_params (nshells) 1-based; momentum, nprim (nparams); cntr, exps
 * (max_nprim, nparams) with primitive-normalized coefficients; norm
 * (max_nang, nparams) = NormFactorsSpher if spher else NormFactorsCart. */
int gi_set_basis(gi_context *ctx, int64_t natoms, const double *coords, int64_t nshells,
                 const int64_t *shell_centers, const int64_t *shell_params, int64_t nparams,
                 const int64_t *momentum, const int64_t *nprim, int64_t max_nprim,
                 const double *cntr, const double *exps, int64_t max_nang, const double *norm,
                 int64_t spher);

/* Shell-quartet ERI blocks, the operation behind the reference's
 * Auto2eERI(idx)%ptr(...) frontends (auto2e.f90:271-2880).  quartets
 * (4, nq) holds 0-based shells (A,B,C,D); block q is written at
 * out + offsets[q] as V(a,b,c,d) = (ab|cd), a fastest.  Host pointers. */
int gi_eri_quartets(gi_context *ctx, int64_t nq, const int64_t *quartets, double kappa,
                    const int64_t *offsets, double *out, int64_t out_len);

/* Dense block of (pq|rs) between two lists of shell pairs, the product form
 * of the ERI launches behind chol2_M / chol2_Vabrs (TwoStepCholesky_Step1.f90:733-856,
 * TwoStepCholesky_Step2.f90:242-363) without screening or selection: row pair
 * i = (row_pairs[2i], row_pairs[2i+1]) (0-based shells) contributes its
 * n0*n1 orbital pairs (a fastest) as consecutive rows, column pairs likewise
 * as consecutive columns; out is column-major with leading dimension ldo.
 * Host pointers. */
int gi_eri_pair_block(gi_context *ctx, int64_t nrow_pairs, const int64_t *row_pairs, int64_t ncol_pairs,
                      const int64_t *col_pairs, double kappa, double *out, int64_t ldo);

/* Tuning options of the ERI launch layer (no reference counterpart):
 *   "eri_min_tiles"  keep at least this many work tiles per MFMA launch (8192)
 *   "eri_max_kt"     most ket pairs stacked in one MFMA tile (32)
 *   "eri_lane"       1: small uncontracted classes on the lane-per-quartet
 *                    kernels, 0: every class on the MFMA kernels (1)
 *   "eri_trans"      1: MFMA launches put the side with more rows per pair
 *                    on the 16-row tiles, 0: never transpose (1)
 *   "eri_lane_emax"  lane kernels only for classes with E, F <= this (4)
 *   "vec_blocks"     block rows of the triangular Inv_L Vabrs^T product of
 *                    the full-dimension vectors (20; at most NVecs / 256, each
 *                    a multiple of 128 rows counted from the last row)
 *   "thc_schur_mb"   gi_thc_grid_pivots keeps the Schur complement of S over
 *                    the candidates explicitly once it fits this many MiB
 *                    (24576; 0: columns built on the fly throughout)
 *   "thc_cholesky"   gi_thc_z factors S by Cholesky first, LU when a pivot
 *                    is not positive (1); 0: LU always
 * Unknown names fail.  Results do not depend on these settings beyond
 * floating-point rounding. */
int gi_set_option(gi_context *ctx, const char *name, int64_t value);

/* Two-step Cholesky decomposition of the Coulomb matrix, the work of
 * chol_gammcor_Rkpq -> chol_gammcor_CoulombMatrix (Cholesky_Gammcor.f90:335-441):
 * Step 1 (pivots), Step 2 (Inv_L) and the full-dimension vectors Rkpq, all
 * kept on the device.  tau: CholeskyTauThresh; kappa = 1/Omega^2 or 0;
 * max_block_dim = MaxBlockDim (4000), vec_block = CholVecsBlock (500)
 * (TwoStepCholesky_definitions.f90:20-53). */
int gi_chol_build(gi_context *ctx, double tau, double kappa, int64_t max_block_dim,
                  int64_t vec_block, int64_t flags);
/* Batched form (no single reference counterpart; the reference decomposes one
 * system per chol_gammcor_Rkpq call): the decompositions of n independent
 * systems, one per context (each with its basis set), run concurrently from
 * `nworkers` host threads, each context on its own stream, so that small
 * systems overlap on the GPU.  Same arguments and results per context as
 * gi_chol_build; status[i] receives each entry's status, and the call fails
 * (naming the first failed entry) if any entry failed. */
int gi_chol_build_batch(int64_t n, gi_context *const *ctxs, double tau, double kappa, int64_t max_block_dim,
                        int64_t vec_block, int64_t flags, int64_t nworkers, int64_t *status);
/* Differentiation form (no reference counterpart: the reference has no
 * gradient code).  The Coulomb energy of a symmetric AO density,
 *   E_J = 1/2 sum_{pqrs} D_pq (pq|rs) D_rs,
 * D column-major (nao x nao, leading dimension ldd >= nao, the AO order of
 * gi_set_basis), kappa as in gi_chol_build (0: plain Coulomb), and, when grad
 * is not NULL, dE_J/dR for every atom (3 x natoms, x fastest, Eh/bohr) from
 * derivative integrals evaluated on the fly.  The gradient needs l0 + l1 <= 5
 * for every shell pair. */
int gi_coulomb_energy_grad(gi_context *ctx, const double *density, int64_t ldd, double kappa, double *energy,
                           double *grad);
/* One-electron integrals in the AO basis of the context (column-major,
 * leading dimension ldh >= nao, both triangles filled): the overlap S
 * (ints1e_S, OneElectronInts.f90:915-932), the kinetic energy T (ints1e_T,
 * :935-953) and the nuclear attraction V (ints1e_Vne, :956-975), any of them
 * NULL to skip.  atom_charges[natoms] are the nuclear charges the reference
 * takes from TSystem (ZNumbers or ZNumbersECP, 0 for dummy atoms; NULL for
 * none), and pc_coords (3 x npc, bohr) / pc_charges (npc) the external point
 * charges (TSystem%PointChargeCoords / PointCharges).  Shells up to l = 3. */
int gi_one_electron(gi_context *ctx, const double *atom_charges, int64_t npc, const double *pc_coords,
                    const double *pc_charges, double *S, double *T, double *V, int64_t ldh);
/* The bare-nucleus Hamiltonian H0 = T + V_ne in the AO basis
 * (ints1e_gammcor_H0, OneElectronInts_Gammcor.f90:13-49); arguments as above. */
int gi_h0(gi_context *ctx, const double *atom_charges, int64_t npc, const double *pc_coords,
          const double *pc_charges, double *H0, int64_t ldh);
/* tau of a CHOL_ACCURACY_* level; returns 0 and sets *tau, or fails with
 * "Invalid accuracy setting" (Cholesky_Gammcor.f90:407-409). */
int gi_chol_accuracy_tau(int64_t accuracy, double *tau);

/* Sizes of the TChol2Vecs components (TwoStepCholesky_definitions.f90:56-74):
 * dims[0] NVecs, [1] NOrbPairs, [2] NShellPairs, [3] NSubsets(1),
 * [4] MaxSubsetDim, [5] NPivotShellPairs, [6] macro iterations,
 * [7] shell quartets evaluated. */
int gi_chol_dims(gi_context *ctx, int64_t dims[8]);
/* Fill caller-allocated TChol2Vecs arrays (any pointer may be NULL):
 * ShellPairs(2,NShellPairs), ShellPairLoc(3,NShellPairs), ShellPairDim,
 * SubsetDim, SubsetBounds(2,NSubsets), Pivots(NOrbPairs),
 * pivot_order(NVecs) (FULL-storage locations in pivot order),
 * PivotShellPairs, PivotShellPairLoc, PivotShellPairDim (NPivotShellPairs),
 * PivotOrbPairs(NVecs). */
int gi_chol_meta(gi_context *ctx, int64_t *shell_pairs, int64_t *shell_pair_loc,
                 int64_t *shell_pair_dim, int64_t *subset_dim, int64_t *subset_bounds,
                 int64_t *pivots, int64_t *pivot_order, int64_t *pivot_shell_pairs,
                 int64_t *pivot_shell_pair_loc, int64_t *pivot_shell_pair_dim,
                 int64_t *pivot_orb_pairs);
/* Rkpq(NVecs, MaxSubsetDim, NSubsets) (TCholeskyVecsOTF%Rkpq,
 * Cholesky_Gammcor.f90:19-22) copied to host memory. */
int gi_chol_rkpq(gi_context *ctx, double *dst);
/* Device address of Rkpq (valid until the next gi_chol_build / gi_destroy). */
int gi_chol_rkpq_device(gi_context *ctx, double **ptr);
/* Inv_L(NVecs, NVecs) and Vdiag(NAO(NAO+1)/2) to host memory. */
int gi_chol_inv_l(gi_context *ctx, double *dst);
int gi_chol_vdiag(gi_context *ctx, double *dst);
/* Reconstruction error of the last build on its diagonal: over the BASE
 * orbital pairs pq, r(pq) = (pq|pq) - sum_k Rkpq(k, pq)^2, the quantity the
 * decomposition bounds by tau (TwoStepCholesky_Step1.f90:1004-1064 stops when
 * the residual diagonal is below tau).  Returns its maximum and minimum. */
int gi_chol_residual(gi_context *ctx, double *max_residual, double *min_residual);
/* Phase timings in seconds: [0] step 1, [1] step 2, [2] full vectors,
 * [3] total build, [4] ERI kernel seconds (GI_CHOL_TIME_ERI), [5] ERI FP64
 * operations, [6] Vdiag, [7] step-1 ERIs, [8] M - R^T R, [9] micro
 * iterations, [10] V(P,P) ERIs, [11] POTRF+TRTRI, [12] Vabrs ERIs,
 * [13] Inv_L V^T GEMM, [14] ERI kernel launches (count), [15] M - R^T R
 * FP64 operations, [16] host work of the candidate definition and
 * compression, [17] host work of the batch definition (both in step 1). */
int gi_chol_timings(gi_context *ctx, double t[18]);

/* Per-kernel ERI statistics (diagnostic; no reference counterpart), gathered
 * when the environment variable GIGPU_ERI_STATS is set: up to `cap` rows of
 * (E = la+lb, F = lc+ld, quartets, kernel milliseconds, FP64 operations),
 * accumulated since the context was created.  *n = rows written. */
int gi_eri_class_stats(gi_context *ctx, int64_t cap, double *rows, int64_t *n);

/* AO -> MO transform of the Cholesky vectors, chol_gammcor_MOTransf_2
 * (Cholesky_Gammcor.f90:143-332): Rkab(NVecs, NA*NB), a fastest,
 * Rkab(k, a, b) = sum_pq R_k(p,q) CA(p,a) CB(q,b), a = a0..a1, b = b0..b1
 * (1-based columns of CA (nao, *) ld lda and CB ld ldb, rows already in the
 * Auto2e AO order as chol_gammcor_Rkab passes them, :74-83).  Fails with
 * "Buffer size too small for Rpq->Rab" like :176-180.  Host pointers. */
int gi_mo_transform(gi_context *ctx, const double *CA, int64_t lda, int64_t a0, int64_t a1,
                    const double *CB, int64_t ldb, int64_t b0, int64_t b1,
                    int64_t max_buffer_mb, int64_t flags, double *Rkab);
/* Same with device pointers for CA, CB and Rkab. */
int gi_mo_transform_device(gi_context *ctx, const double *CA, int64_t lda, int64_t a0,
                           int64_t a1, const double *CB, int64_t ldb, int64_t b0, int64_t b1,
                           int64_t max_buffer_mb, int64_t flags, double *Rkab);

/* External-program AO orderings (ORBITAL_ORDERING_*, Auto2eInterface.f90:12-17). */
#define GI_ORDERING_AUTO2E 0
#define GI_ORDERING_MOLPRO 1
#define GI_ORDERING_ORCA 2
#define GI_ORDERING_DALTON 3
#define GI_ORDERING_OPEN_MOLCAS 4
#define GI_ORDERING_PYSCF 5

/* auto2e_interface_AOMap (Auto2eInterface.f90:378-514) for the context's
 * basis: map[p] (1-based) = Auto2e position of the external program's AO p
 * (Map_extao2ao), nao entries.  phase[q] (Auto2e order) = -1 for the Orca
 * sign convention of spherical f, g, h functions with m = +-3 and g, h with
 * m = +-4 (auto2e_interface_ApplyOrcaPhases_Vector, :517-562), +1 otherwise.
 * Fails like the reference's "Invalid external orbital ordering" (:439,
 * :488) for GI_ORDERING_AUTO2E and for Cartesian AOs in any ordering
 * but Molpro's. */
int gi_ao_map(gi_context *ctx, int64_t ordering, int64_t *map, double *phase);
/* The rows of a column-major matrix of MO coefficients X_ext (nao x ncols,
 * ld_ext) from the external AO order to the Auto2e order with the Orca signs
 * applied, on the device: X_ao(map[p]-1, j) = phase(map[p]-1) X_ext(p, j) --
 * auto2e_interface_AngFuncTransf + ApplyOrcaPhases_Matrix as
 * chol_gammcor_Rkab applies them to CA and CB (Cholesky_Gammcor.f90:73-83).
 * Device pointers, X_ao distinct from X_ext; the result can go straight to
 * gi_mo_transform_device. */
int gi_ao_rows_to_auto2e_device(gi_context *ctx, int64_t ordering, const double *X_ext, int64_t ld_ext,
                                int64_t ncols, double *X_ao, int64_t ld_ao);
/* Allocation-free variant of gi_ao_rows_to_auto2e_device (no counterpart in
 * the reference; for callers that cannot allocate device memory: code under
 * HIP graph capture, tight loops that forbid hipMalloc).  Same arguments and
 * result; the map and signs are staged in a 1 MiB device array that the
 * library holds per device for the life of the process instead of device
 * allocations, so the basis may have at most 65536 AOs (larger fails).  Calls
 * that overlap on one device race for that array. */
int gi_ao_rows_to_auto2e_device_noalloc(gi_context *ctx, int64_t ordering, const double *X_ext, int64_t ld_ext,
                                        int64_t ncols, double *X_ao, int64_t ld_ao);

/* ---- tensor hypercontraction (THC) of the Cholesky-decomposed Coulomb
 * matrix, the second consumer of the decomposition
 * (thc_CoulombMatrix_QuadraticMemory, TensorHypercontraction.f90:232-272;
 * thc_gammcor_Rkab, THC_Gammcor.f90:25-80).  The grid points are the
 * caller's: the Becke parent grid and its pruning (thc_Grid,
 * TensorHypercontraction.f90:300-390) are not part of this library. */

/* Pruning of a parent grid by the pivots of S(g,h) = ((Xgp Xgp^T)(g,h))^2,
 * thc_Chol_Pivots (TensorHypercontraction.f90:772-968) with thc_Chol_NextBlock
 * (:971-1064) and thc_Chol_Compress (:1067-1127): a pivoted Cholesky of S with
 * its columns built on the fly from the row-normalized orbitals, in blocks of
 * at most block_dim pivots (THC_BlockDim, 500), the points whose residual
 * diagonal falls to PivotThresh = qr_thresh^2 dropped after every block.
 * xyz (3, npts) bohr, host.  On return pivots[0 .. *npivots - 1] hold the
 * 0-based indices of the chosen points in pivot order (the reference returns
 * their coordinates in that order) and *npivots_reduced the count of leading
 * pivots whose diagonal exceeded qr_thresh_reduced^2 (NGridReduced; equal to
 * *npivots when qr_thresh_reduced <= 0).  pivots holds npts entries.  Fails
 * like the reference ("Cholesky decomposistion failed to converge") when
 * ceil(npts / block_dim) blocks leave candidates above the threshold, and
 * when every orbital vanishes at a point. */
int gi_thc_grid_pivots(gi_context *ctx, int64_t npts, const double *xyz, double qr_thresh,
                       double qr_thresh_reduced, int64_t block_dim, int64_t *npivots, int64_t *npivots_reduced,
                       int64_t *pivots);

/* Atomic orbitals on grid points, gridfunc_Orbitals (GridFunctions.f90:154-240)
 * in the Auto2e AO order of the basis, followed when normalize != 0 by
 * thc_normalize_Xgp (TensorHypercontraction.f90:600-637): every row divided
 * by its 2-norm.  xyz (3, npts) bohr; Xgp (npts, nao) column-major, leading
 * dimension ldx.  Fails when a normalized row is zero (the reference assumes
 * the grid was screened so that it never is).  Host pointers. */
int gi_thc_orbitals(gi_context *ctx, int64_t npts, const double *xyz, int64_t normalize, double *Xgp,
                    int64_t ldx);
/* Same with device pointers for xyz and Xgp. */
int gi_thc_orbitals_device(gi_context *ctx, int64_t npts, const double *xyz, int64_t normalize, double *Xgp,
                           int64_t ldx);

/* Least-squares THC factor of the context's Cholesky vectors, thc_Z
 * (TensorHypercontraction.f90:393-476) with the precomputed vectors
 * (QuadraticMemory = .false.): Zgk (npts, NVecs) solves S Z = XR with
 * XR = 2 Xgpq R^T over the stored orbital pairs, diagonal products halved
 * (thc_XR, :479-577; thc_Xgpq, :640-769), and S(g,h) = ((Xgp Xgp^T)(g,h))^2
 * (thc_S, :580-597).  When 0 < npts_reduced < npts, Zgk_reduced
 * (npts_reduced, NVecs) solves the same system on the first npts_reduced
 * points (:455-466); otherwise it is not written.  Needs gi_chol_build
 * first.  S is factored by Cholesky, or by LU with partial pivoting where
 * a Cholesky pivot is not positive (the reference's dsysv is the symmetric
 * indefinite solver); fails when S is exactly singular.  Host pointers. */
int gi_thc_z(gi_context *ctx, int64_t npts, const double *Xgp, int64_t ldx, int64_t npts_reduced, double *Zgk,
             int64_t ldz, double *Zgk_reduced, int64_t ldzr);
/* Same with device pointers for Xgp, Zgk and Zgk_reduced. */
int gi_thc_z_device(gi_context *ctx, int64_t npts, const double *Xgp, int64_t ldx, int64_t npts_reduced,
                    double *Zgk, int64_t ldz, double *Zgk_reduced, int64_t ldzr);

/* MO-basis THC vectors, the tail of thc_gammcor_Rkab (THC_Gammcor.f90:69-79):
 * Xga = Xgp C_ao(:, a0:a1) (thc_gammcor_Xga, :175-191, C_ao = the rows of CA
 * taken from the external AO order `ordering` to the Auto2e order with the
 * Orca signs; GI_ORDERING_AUTO2E takes the rows as they are), likewise Xgb,
 * and Rkab(k, a, b) = sum_g Zgk(g, k) Xga(g, a) Xgb(g, b), a fastest, nvec
 * rows (thc_gammcor_Rkab_2, :82-106).  CA (nao, *) ld lda, CB ld ldb,
 * 1-based columns.  Host pointers. */
int gi_thc_rkab(gi_context *ctx, int64_t npts, const double *Xgp, int64_t ldx, const double *Zgk, int64_t ldz,
                int64_t nvec, const double *CA, int64_t lda, int64_t a0, int64_t a1, const double *CB,
                int64_t ldb, int64_t b0, int64_t b1, int64_t ordering, double *Rkab);
/* Same with device pointers for Xgp, Zgk, CA, CB and Rkab. */
int gi_thc_rkab_device(gi_context *ctx, int64_t npts, const double *Xgp, int64_t ldx, const double *Zgk,
                       int64_t ldz, int64_t nvec, const double *CA, int64_t lda, int64_t a0, int64_t a1,
                       const double *CB, int64_t ldb, int64_t b0, int64_t b1, int64_t ordering, double *Rkab);

#ifdef __cplusplus
}
#endif
#endif
