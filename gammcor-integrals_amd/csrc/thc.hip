// Tensor hypercontraction of the Cholesky-decomposed Coulomb matrix on the
// device, the second consumer of the decomposition:
//   orbitals on the grid   gridfunc_Orbitals + thc_normalize_Xgp
//                          (GridFunctions.f90:154-240, TensorHypercontraction.f90:600-637)
//   least-squares Z        thc_Z -> thc_XR, thc_S, real_Axb_symmetric_sysv
//                          (TensorHypercontraction.f90:393-597, real_linalg.f90:656-703)
//   MO-basis vectors       thc_gammcor_Rkab tail: thc_gammcor_Xga, thc_gammcor_Rkab_2
//                          (THC_Gammcor.f90:25-106, 175-191)
// The grid points are the caller's (the Becke/Lebedev parent grid and its
// pruning, thc_Grid, are not part of this library).
//
// Z solves S Z = XR with
//   XR(g, k) = 2 sum_{pq in storage} Xgpq(g, pq) R(k, pq),  Xgpq = X(g,p) X(g,q),
//              halved on the diagonal p = q,
//   S(g, h)  = (sum_p X(g,p) X(h,p))^2.
// XR runs over the stored Rkpq columns in chunks: one kernel writes the
// products of a chunk (zero for padding columns), one dgemm adds
// 2 Xgpq R^T.  S is a dgemm and an elementwise square.  The symmetric
// indefinite solve of the reference (LAPACK dsysv) is replaced by LU with
// partial pivoting (rocsolver_dgesv): both are backward stable, so the two
// solutions agree to cond(S) times the rounding of the arithmetic.
#include <algorithm>
#include <cmath>
#include <rocsolver/rocsolver.h>
#include "gi_context.h"

namespace gi {

#define THC_BLAS(x)                                                                                          \
  do {                                                                                                       \
    rocblas_status s_ = (x);                                                                                 \
    if (s_ != rocblas_status_success)                                                                        \
      throw Error{2, std::string("rocBLAS/rocSOLVER error ") + rocblas_status_to_string(s_) + " at " +      \
                         __FILE__ + ":" + std::to_string(__LINE__)};                                         \
  } while (0)

// device work arrays of the THC calls, grow-only, kept in the context
struct ThcWork {
  DBuf<int> shl, aoloc, flag, ipiv, info;
  DBuf<int> colpq;                    // (p, q) of every stored Rkpq column, -1 for padding
  DBuf<double> W, S, Sr, CA, CB, Xa, Xb, part;
};
static ThcWork &thc_work(Context &C) {
  if (!C.thcw) C.thcw = new ThcWork();
  return *C.thcw;
}
void release_thc_work(Context &C) {
  delete C.thcw;
  C.thcw = nullptr;
}

// ------------------------------------------------------------------ kernels
// thc_normalize_Xgp: X(g, :) /= ||X(g, :)||_2, scaled by the reciprocal like
// the reference.  The AOs are split in nchunk column ranges so that a launch
// has ceil(npts / 256) x nchunk workgroups: k_thc_sumsq writes the partial
// sums of squares part(g, c), k_thc_scale adds them in a fixed order (the
// result does not depend on the launch) and scales its range.  A point where
// every orbital is zero raises *flag and is left as it is.
__global__ void __launch_bounds__(256) k_thc_sumsq(const double *__restrict__ X, long long ldx, long long npts,
                                                   long long nao, long long cw, double *__restrict__ part) {
  const long long g = blockIdx.x * 256LL + threadIdx.x;
  if (g >= npts) return;
  const long long p0 = blockIdx.y * cw, p1 = std::min(nao, p0 + cw);
  double s = 0.0;
  for (long long p = p0; p < p1; ++p) {
    const double v = X[g + ldx * p];
    s += v * v;
  }
  part[g + npts * blockIdx.y] = s;
}

__global__ void __launch_bounds__(256) k_thc_scale(double *__restrict__ X, long long ldx, long long npts,
                                                   long long nao, long long cw, int nchunk,
                                                   const double *__restrict__ part, int *__restrict__ flag) {
  const long long g = blockIdx.x * 256LL + threadIdx.x;
  if (g >= npts) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[g + npts * c];
  if (!(s > 0.0)) {
    if (blockIdx.y == 0) *flag = 1;
    return;
  }
  const double r = 1.0 / sqrt(s);
  const long long p0 = blockIdx.y * cw, p1 = std::min(nao, p0 + cw);
  for (long long p = p0; p < p1; ++p) X[g + ldx * p] *= r;
}

// Xgpq(g, c) for the stored columns c0 .. c0 + nc - 1 (thc_Xgpq,
// TensorHypercontraction.f90:640-769): X(g,p) X(g,q), times 1/2 when p = q,
// 0 for padding; W column-major with leading dimension npts.
__global__ void __launch_bounds__(256) k_thc_xgpq(const double *__restrict__ X, long long ldx, long long npts,
                                                  const int *__restrict__ colpq, long long c0, long long nc,
                                                  double *__restrict__ W) {
  const long long g = blockIdx.x * 256LL + threadIdx.x;
  if (g >= npts) return;
  for (long long c = blockIdx.y; c < nc; c += gridDim.y) {
    const int p = colpq[2 * (c0 + c)], q = colpq[2 * (c0 + c) + 1];
    double v = 0.0;
    if (p >= 0) {
      v = X[g + ldx * p] * X[g + ldx * q];
      if (p == q) v *= 0.5;
    }
    W[g + npts * c] = v;
  }
}

// S(g, h) = S(g, h)^2 on an n x n matrix (thc_S)
__global__ void __launch_bounds__(256) k_thc_square(double *__restrict__ S, long long lds, long long n) {
  const long long e = blockIdx.x * 256LL + threadIdx.x;
  if (e >= n * n) return;
  const long long g = e % n, h = e / n;
  const double v = S[g + lds * h];
  S[g + lds * h] = v * v;
}

// W(g, a + na j) = Xa(g, a) Xb(g, j) for j < nbc: the point-wise orbital
// products that thc_gammcor_Rkab_2 contracts with Z
__global__ void __launch_bounds__(256) k_thc_pair_products(const double *__restrict__ Xa, const double *__restrict__ Xb,
                                                           long long npts, long long na, long long nbc,
                                                           double *__restrict__ W) {
  const long long g = blockIdx.x * 256LL + threadIdx.x;
  if (g >= npts) return;
  for (long long c = blockIdx.y; c < na * nbc; c += gridDim.y) {
    const long long a = c % na, j = c / na;
    W[g + npts * c] = Xa[g + npts * a] * Xb[g + npts * j];
  }
}

// ------------------------------------------------------------------ host
static unsigned ygrid(long long n) { return (unsigned)std::min<long long>(std::max<long long>(n, 1), 65535); }

void thc_orbitals(Context &C, const double *dxyz, long long npts, int normalize, double *dX, long long ldx) {
  if (!C.basis_set) throw Error{4, "basis not set"};
  if (npts < 0) throw Error{81, "invalid number of grid points"};
  if (ldx < std::max<long long>(npts, 1)) throw Error{81, "leading dimension of Xgp smaller than the number of points"};
  if (npts == 0) return;
  ThcWork &w = thc_work(C);
  std::vector<int> shl, off(MAXL + 2, 0);
  for (int L = 0; L <= MAXL; ++L) {
    for (int s = 0; s < C.nshells; ++s)
      if (C.shell_l[s] == L) shl.push_back(s);
    off[L + 1] = (int)shl.size();
  }
  if (off[MAXL + 1] != C.nshells) throw Error{80, "orbitals on grid points need l <= 5"};
  w.shl.upload(shl, C.stream);
  w.aoloc.upload(std::vector<int>(C.shell_loc.begin(), C.shell_loc.end()), C.stream);
  for (int L = 0; L <= MAXL; ++L)
    launch_thc_orbitals(C, L, w.aoloc.p, w.shl.p + off[L], off[L + 1] - off[L], dxyz, npts, dX, ldx);
  if (normalize) {
    const long long nao = C.nao, cw = std::max<long long>(16, (nao + 63) / 64);
    const int nchunk = (int)((nao + cw - 1) / cw);
    w.flag.alloc(1);
    w.part.alloc((size_t)(npts * nchunk));
    GI_HIP(hipMemsetAsync(w.flag.p, 0, sizeof(int), C.stream));
    const dim3 gn((unsigned)((npts + 255) / 256), (unsigned)nchunk);
    hipLaunchKernelGGL(k_thc_sumsq, gn, dim3(256), 0, C.stream, dX, ldx, npts, nao, cw, w.part.p);
    hipLaunchKernelGGL(k_thc_scale, gn, dim3(256), 0, C.stream, dX, ldx, npts, nao, cw, nchunk, w.part.p, w.flag.p);
    GI_HIP(hipGetLastError());
    int bad = 0;
    C.read(&bad, w.flag.p, sizeof(int));
    if (bad) throw Error{82, "every orbital vanishes at a grid point: Xgp cannot be normalized"};
  }
}

// (p, q) of each stored Rkpq column, in the order of the Cholesky storage:
// for shell pair (A, B) at storage position pos0 of subset X, b = 0..nb-1
// outer, a = (A == B ? b : 0)..na-1 inner (thc_Xgpq_ab / thc_Xgpq_aa)
static void thc_column_pairs(const Context &C, std::vector<int> &colpq) {
  const CholState &S = C.chol;
  const long long msd = S.max_subset_dim;
  colpq.assign((size_t)(2 * msd * S.nsubsets), -1);
  for (long long s = 0; s < S.nshell_pairs; ++s) {
    const int A = (int)S.shell_pairs[2 * s] - 1, B = (int)S.shell_pairs[2 * s + 1] - 1;
    const long long X = S.shell_pair_loc[3 * s + 2], pos0 = S.shell_pair_loc[3 * s + 1] - 1;
    const int na = C.shell_nang[A], nb = C.shell_nang[B];
    long long qq = 0;
    for (int b = 0; b < nb; ++b)
      for (int a = (A == B ? b : 0); a < na; ++a) {
        const long long col = (X - 1) * msd + pos0 + qq++;
        colpq[2 * col] = C.shell_loc[A] + a;
        colpq[2 * col + 1] = C.shell_loc[B] + b;
      }
  }
}

// solve S Z = B in place of B (n x nrhs, ld ldb); S (n x n, ld n) is overwritten
static void thc_solve(Context &C, double *S, long long n, double *B, long long ldb, long long nrhs) {
  ThcWork &w = thc_work(C);
  w.ipiv.alloc((size_t)n);
  w.info.alloc(1);
  THC_BLAS(rocsolver_dgesv(C.blas, (rocblas_int)n, (rocblas_int)nrhs, S, (rocblas_int)n, w.ipiv.p, B,
                           (rocblas_int)ldb, w.info.p));
  int info = 0;
  C.read(&info, w.info.p, sizeof(int));
  if (info != 0)
    throw Error{83, "THC metric matrix S is singular (LU pivot " + std::to_string(info) + " is zero)"};
}

void thc_z(Context &C, const double *dX, long long ldx, long long npts, long long nred, double *dZ, long long ldz,
           double *dZr, long long ldzr) {
  CholState &S = C.chol;
  if (!S.valid || !S.rkpq.p) throw Error{50, "Cholesky vectors are not available (run gi_chol_build first)"};
  const long long nvec = S.nvecs, nao = C.nao, msd = S.max_subset_dim, ncols = msd * S.nsubsets;
  if (npts <= 0) throw Error{81, "invalid number of grid points"};
  if (ldx < npts || ldz < npts) throw Error{81, "leading dimension smaller than the number of points"};
  const bool reduced = nred > 0 && nred < npts;
  if (reduced && (!dZr || ldzr < nred)) throw Error{81, "invalid ZgkReduced"};
  if (npts > (1LL << 31) - 1 || nvec > (1LL << 31) - 1) throw Error{81, "grid too large for the LU solver"};
  ThcWork &w = thc_work(C);
  std::vector<int> colpq;
  thc_column_pairs(C, colpq);
  w.colpq.upload(colpq, C.stream);
  // XR = 2 Xgpq R^T in column chunks of at most 2^26 doubles of Xgpq
  const long long nc = std::max<long long>(1, std::min<long long>(ncols, (1LL << 26) / npts));
  w.W.alloc((size_t)(npts * nc));
  const double two = 2.0, one = 1.0, zero = 0.0;
  const dim3 gp((unsigned)((npts + 255) / 256), ygrid(nc));
  if (ncols == 0 || nvec == 0)
    GI_HIP(hipMemset2DAsync(dZ, sizeof(double) * ldz, 0, sizeof(double) * npts, nvec, C.stream));
  for (long long c0 = 0; c0 < ncols && nvec > 0; c0 += nc) {
    const long long m = std::min(nc, ncols - c0);
    hipLaunchKernelGGL(k_thc_xgpq, gp, dim3(256), 0, C.stream, dX, ldx, npts, w.colpq.p, c0, m, w.W.p);
    GI_HIP(hipGetLastError());
    THC_BLAS(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_transpose, npts, nvec, m, &two, w.W.p,
                           npts, S.rkpq.p + c0 * nvec, nvec, c0 == 0 ? &zero : &one, dZ, ldz));
  }
  if (reduced && nvec > 0)
    GI_HIP(hipMemcpy2DAsync(dZr, sizeof(double) * ldzr, dZ, sizeof(double) * ldz, sizeof(double) * nred, nvec,
                            hipMemcpyDeviceToDevice, C.stream));
  // S = (X X^T)^2, its leading block for the reduced grid
  w.S.alloc((size_t)(npts * npts));
  THC_BLAS(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_transpose, npts, npts, nao, &one, dX, ldx,
                         dX, ldx, &zero, w.S.p, npts));
  hipLaunchKernelGGL(k_thc_square, dim3((unsigned)((npts * npts + 255) / 256)), dim3(256), 0, C.stream, w.S.p, npts,
                     npts);
  GI_HIP(hipGetLastError());
  if (reduced) {
    w.Sr.alloc((size_t)(nred * nred));
    GI_HIP(hipMemcpy2DAsync(w.Sr.p, sizeof(double) * nred, w.S.p, sizeof(double) * npts, sizeof(double) * nred, nred,
                            hipMemcpyDeviceToDevice, C.stream));
  }
  if (nvec > 0) {
    thc_solve(C, w.S.p, npts, dZ, ldz, nvec);
    if (reduced) thc_solve(C, w.Sr.p, nred, dZr, ldzr, nvec);
  }
  C.wait();
}

void thc_rkab(Context &C, const double *dX, long long ldx, long long npts, const double *dZ, long long ldz,
              long long nvec, const double *dCA, long long lda, long long a0, long long a1, const double *dCB,
              long long ldb, long long b0, long long b1, int ordering, double *dRkab) {
  if (!C.basis_set) throw Error{4, "basis not set"};
  const long long nao = C.nao, NA = a1 - a0 + 1, NB = b1 - b0 + 1;
  if (NA <= 0 || NB <= 0 || a0 < 1 || b0 < 1) throw Error{51, "Invalid dimension of Rkab"};
  if (lda < nao || ldb < nao) throw Error{52, "Invalid number of atomic orbitals in CA/CB"};
  if (npts <= 0 || nvec <= 0 || ldx < npts || ldz < npts) throw Error{81, "invalid Xgp or Zgk"};
  ThcWork &w = thc_work(C);
  // thc_gammcor_Xga: C_ao = auto2e_interface_C(C_ext), Xga = Xgp C_ao
  const double *CAo = dCA + (a0 - 1) * lda, *CBo = dCB + (b0 - 1) * ldb;
  long long lda_o = lda, ldb_o = ldb;
  if (ordering != 0) {
    w.CA.alloc((size_t)(nao * NA));
    w.CB.alloc((size_t)(nao * NB));
    ao_rows_to_auto2e(C, ordering, CAo, lda, NA, w.CA.p, nao);
    ao_rows_to_auto2e(C, ordering, CBo, ldb, NB, w.CB.p, nao);
    CAo = w.CA.p;
    CBo = w.CB.p;
    lda_o = ldb_o = nao;
  }
  w.Xa.alloc((size_t)(npts * NA));
  w.Xb.alloc((size_t)(npts * NB));
  const double one = 1.0, zero = 0.0;
  THC_BLAS(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_none, npts, NA, nao, &one, dX, ldx, CAo,
                         lda_o, &zero, w.Xa.p, npts));
  THC_BLAS(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_none, npts, NB, nao, &one, dX, ldx, CBo,
                         ldb_o, &zero, w.Xb.p, npts));
  // Rkab(k, a, b) = sum_g Z(g, k) Xa(g, a) Xb(g, b), b in chunks of at most
  // 2^26 doubles of products
  const long long nbc = std::max<long long>(1, std::min<long long>(NB, (1LL << 26) / (npts * NA)));
  w.W.alloc((size_t)(npts * NA * nbc));
  for (long long b = 0; b < NB; b += nbc) {
    const long long m = std::min(nbc, NB - b);
    const dim3 gp((unsigned)((npts + 255) / 256), ygrid(NA * m));
    hipLaunchKernelGGL(k_thc_pair_products, gp, dim3(256), 0, C.stream, w.Xa.p, w.Xb.p + npts * b, npts, NA, m,
                       w.W.p);
    GI_HIP(hipGetLastError());
    THC_BLAS(rocblas_dgemm(C.blas, rocblas_operation_transpose, rocblas_operation_none, nvec, NA * m, npts, &one, dZ,
                           ldz, w.W.p, npts, &zero, dRkab + nvec * NA * b, nvec));
  }
  C.wait();
}

}  // namespace gi
