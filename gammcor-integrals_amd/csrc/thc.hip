// Tensor hypercontraction of the Cholesky-decomposed Coulomb matrix on the
// device, the second consumer of the decomposition:
//   orbitals on the grid   gridfunc_Orbitals + thc_normalize_Xgp
//                          (GridFunctions.f90:154-240, TensorHypercontraction.f90:600-637)
//   least-squares Z        thc_Z -> thc_XR, thc_S, real_Axb_symmetric_sysv
//                          (TensorHypercontraction.f90:393-597, real_linalg.f90:656-703)
//   MO-basis vectors       thc_gammcor_Rkab tail: thc_gammcor_Xga, thc_gammcor_Rkab_2
//                          (THC_Gammcor.f90:25-106, 175-191)
//   grid pruning           thc_Chol_Pivots, thc_Chol_NextBlock, thc_Chol_Compress
//                          (TensorHypercontraction.f90:772-1127)
// The parent grid points are the caller's (the Becke/Lebedev grid of
// thc_Grid is not part of this library).
//
// Z solves S Z = XR with
//   XR(g, k) = 2 sum_{pq in storage} Xgpq(g, pq) R(k, pq),  Xgpq = X(g,p) X(g,q),
//              halved on the diagonal p = q,
//   S(g, h)  = (sum_p X(g,p) X(h,p))^2.
// XR runs over the stored Rkpq columns in chunks: one kernel writes the
// products of a chunk (zero for padding columns), one dgemm adds
// 2 Xgpq R^T.  S is a dgemm and an elementwise square.  The reference solves
// with the symmetric indefinite LAPACK dsysv.  S is positive semidefinite (the
// elementwise square of a Gram matrix, Schur product theorem), so the library
// first factors it by the pivot-block Cholesky of the decomposition (factor
// and inverse, then Z = L^-T (L^-1 XR), two dgemms; the reduced grid's factor
// is the leading block of the same one), and solves by LU with partial
// pivoting (rocsolver_dgesv) when a pivot is not positive.  The solutions
// agree with the reference's to cond(S) times the rounding of the arithmetic.
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
  DBuf<double> W, S, Sr, CA, CB, Xa, Xb, part, Li, Lw;
  // grid pruning: candidate coordinates and orbitals, residual diagonal (two
  // buffers, read by a step from one and written to the other), vectors and
  // Schur complements over the candidates (two buffers each: compression
  // gathers into the other), step state, pivots in order
  DBuf<double> pxyz, pX, pD, pD2, pL[2], pS[2];
  DBuf<long long> pcand, pmap, ppiv;
  DBuf<unsigned char> pstate;
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

// ---- grid pruning (pivoted Cholesky of S = (X X^T)^2, thc_Chol_Pivots)
// Two ways to form a step's column over the candidates, with the same pivots
// up to rounding:
//   on the fly   (X X_h^T)^2 - sum_{j<k} L(:,j) L(h,j) over every earlier vector,
//                as thc_Chol_NextBlock does: reads X and all of L per pivot;
//   Schur        Shat(:,h) - sum over the block's vectors only, where Shat is
//                S less the earlier blocks' vectors, kept explicitly (one DGEMM
//                per block updates it): reads one column and the block per pivot.
// The library uses the Schur form when the candidates' Shat fits the
// thc_schur_mb budget and the on-the-fly form before that.
struct PivotState {
  long long npiv;   // pivots taken so far (all blocks)
  long long cur;    // vector index of the step's pivot
  long long nred;   // NPivotsReduced
  long long h;      // candidate index of the step's pivot
  double dh;        // its residual diagonal
  int done;         // no candidate above the threshold: every later step is skipped
  int pad;
};

// One pivot step, launched with vector index k.  Every workgroup finds the
// step's pivot itself, the first candidate index of the largest D (maxloc), so
// that no workgroup waits for another; it is taken when D > thresh
// (thc_Chol_NextBlock, TensorHypercontraction.f90:1012-1027), and workgroup 0
// records it.  Then the vector over every candidate g (column k - lbase of L)
//   on the fly (Sh == null): ((sum_p X(g,p) X(h,p))^2 - sum_j L(g,j) L(h,j)) / sqrt(D(h))
//   Schur:                    (Sh(g,h) - sum_j L(g,j) L(h,j)) / sqrt(D(h))
// and Dout(g) = Din(g) - L(g,k)^2: D is read from one buffer and written to the
// other, so that the workgroups still searching see the step's input.  64
// candidates per workgroup; the sixteen wavefronts split the sums.  A step
// without a pivot sets done, and every later step of the block exits at once.
__global__ void __launch_bounds__(1024) k_thc_pivot_step(const double *__restrict__ X, long long ncand,
                                                         long long nao, const double *__restrict__ Sh,
                                                         double *__restrict__ L, long long lbase, long long k,
                                                         const double *__restrict__ Din, double *__restrict__ Dout,
                                                         const long long *__restrict__ cand,
                                                         long long *__restrict__ piv, PivotState *st, double thresh,
                                                         double thresh_red) {
  __shared__ double s_v[1024];
  __shared__ long long s_i[1024];
  __shared__ double s1[16][64], s2[16][64];
  const int t = threadIdx.x;
  if (st->done) return;   // uniform: set by an earlier launch
  double bv = -1.0;
  long long bi = -1;
  for (long long g = t; g < ncand; g += 1024) {
    const double v = Din[g];
    if (bi < 0 || v > bv) {   // strictly larger: the first index wins within a thread
      bv = v;
      bi = g;
    }
  }
  s_v[t] = bv;
  s_i[t] = bi;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (t < off) {
      const double v2 = s_v[t + off];
      const long long i2 = s_i[t + off];
      if (i2 >= 0 && (s_i[t] < 0 || v2 > s_v[t] || (v2 == s_v[t] && i2 < s_i[t]))) {
        s_v[t] = v2;
        s_i[t] = i2;
      }
    }
    __syncthreads();
  }
  const double dh = s_v[0];
  const long long h = s_i[0];
  if (!(h >= 0 && dh > thresh)) {   // uniform: the same LDS values in every lane
    if (blockIdx.x == 0 && t == 0) st->done = 1;
    return;
  }
  if (blockIdx.x == 0 && t == 0) {
    piv[k] = cand[h];
    st->npiv = k + 1;
    if (dh > thresh_red) st->nred = k + 1;
  }
  const long long kl = k - lbase;
  const int lane = t & 63, w = t >> 6;
  const long long g = blockIdx.x * 64LL + lane;
  double a = 0.0, c = 0.0;
  if (g < ncand) {
    if (!Sh)
      for (long long p = w; p < nao; p += 16) a += X[g + ncand * p] * X[h + ncand * p];
    for (long long j = w; j < kl; j += 16) c += L[g + ncand * j] * L[h + ncand * j];
  }
  s1[w][lane] = a;
  s2[w][lane] = c;
  __syncthreads();
  if (w == 0 && g < ncand) {
    double sg = 0.0, lg = 0.0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      sg += s1[i][lane];
      lg += s2[i][lane];
    }
    const double col = Sh ? Sh[g + ncand * h] : sg * sg;
    const double v = (col - lg) / sqrt(dh);
    L[g + ncand * kl] = v;
    Dout[g] = Din[g] - v * v;
  }
}

// Shat(i, j) of the kept candidates: out(i, j) = in(map[i], map[j])
__global__ void __launch_bounds__(256) k_thc_gather_square(const long long *__restrict__ map, long long nnew,
                                                           long long nold, const double *__restrict__ in,
                                                           double *__restrict__ out) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= nnew) return;
  const long long si = map[i];
  for (long long j = blockIdx.y; j < nnew; j += gridDim.y) out[i + nnew * j] = in[si + nold * map[j]];
}

// compression (thc_Chol_Compress, :1067-1127): candidate rows map[i] of the
// vectors, the residual diagonal, the candidate list and coordinates
__global__ void __launch_bounds__(256) k_thc_compress(const long long *__restrict__ map, long long nnew,
                                                      long long nold, long long nvec,
                                                      const double *__restrict__ Lold, double *__restrict__ Lnew,
                                                      const double *__restrict__ Dold, double *__restrict__ Dnew,
                                                      const long long *__restrict__ cold,
                                                      long long *__restrict__ cnew) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= nnew) return;
  const long long src = map[i];
  for (long long j = blockIdx.y; j < nvec; j += gridDim.y) Lnew[i + nnew * j] = Lold[src + nold * j];
  if (blockIdx.y == 0) {
    Dnew[i] = Dold[src];
    cnew[i] = cold[src];
  }
}

__global__ void k_thc_gather_xyz(const double *__restrict__ xyz, const long long *__restrict__ cand, long long n,
                                 double *__restrict__ out) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= n) return;
  const long long c = cand[i];
  out[3 * i] = xyz[3 * c];
  out[3 * i + 1] = xyz[3 * c + 1];
  out[3 * i + 2] = xyz[3 * c + 2];
}

__global__ void k_thc_fill(double *__restrict__ p, long long n, double v) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_thc_iota(long long *__restrict__ p, long long n) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i < n) p[i] = i;
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
  if (nvec > 0 && C.thc_cholesky) {
    // Cholesky: L over S, L^-1 into Li; Z = Li^T (Li XR) with XR from dZ
    w.Li.alloc((size_t)(npts * npts));
    w.Lw.alloc((size_t)std::max(chol_factor_inverse_scratch(npts), npts * nvec));
    w.info.alloc(1);
    if (chol_factor_inverse(C, w.S.p, w.Li.p, npts, w.Lw.p, w.info.p)) {
      auto apply = [&](long long n, double *Z, long long ld) {
        THC_BLAS(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_none, n, nvec, n, &one, w.Li.p,
                               npts, Z, ld, &zero, w.Lw.p, n));
        THC_BLAS(rocblas_dgemm(C.blas, rocblas_operation_transpose, rocblas_operation_none, n, nvec, n, &one,
                               w.Li.p, npts, w.Lw.p, n, &zero, Z, ld));
      };
      apply(npts, dZ, ldz);
      if (reduced) apply(nred, dZr, ldzr);
      C.wait();
      return;
    }
    // not positive definite: rebuild S (the factorization overwrote it) for LU
    THC_BLAS(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_transpose, npts, npts, nao, &one, dX,
                           ldx, dX, ldx, &zero, w.S.p, npts));
    hipLaunchKernelGGL(k_thc_square, dim3((unsigned)((npts * npts + 255) / 256)), dim3(256), 0, C.stream, w.S.p,
                       npts, npts);
    GI_HIP(hipGetLastError());
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

// thc_Chol_Pivots (TensorHypercontraction.f90:772-968) on the device: blocks of
// at most block_dim pivots, one launch per step (a step after the last pivot
// exits at once), one synchronization per block for the compression.  dxyz: the parent points (3, npts), device.
void thc_grid_pivots(Context &C, const double *dxyz, long long npts, double qr_thresh, double qr_thresh_reduced,
                     long long block_dim, std::vector<long long> &pivots, long long &nred) {
  if (!C.basis_set) throw Error{4, "basis not set"};
  if (npts <= 0) throw Error{81, "invalid number of grid points"};
  if (!(qr_thresh > 0.0)) throw Error{84, "invalid QR threshold"};
  if (block_dim <= 0) throw Error{84, "invalid block dimension"};
  ThcWork &w = thc_work(C);
  const long long nao = C.nao;
  const double thresh = qr_thresh * qr_thresh;
  const double thresh_red = qr_thresh_reduced > 0.0 ? qr_thresh_reduced * qr_thresh_reduced : -1.0;
  const long long max_iters = npts / block_dim + (npts % block_dim > 0);
  const double schur_bytes = (double)C.thc_schur_mb * 1048576.0;
  long long ncand = npts, npiv = 0;
  w.pcand.alloc((size_t)npts);
  w.pmap.alloc((size_t)npts);
  w.pD.alloc((size_t)npts);
  w.pD2.alloc((size_t)npts);
  double *const Dbuf[2] = {w.pD.p, w.pD2.p};
  int dcur = 0;   // the buffer holding the current residual diagonal
  w.ppiv.alloc((size_t)npts);
  w.pstate.alloc(sizeof(PivotState));
  w.pxyz.alloc((size_t)(3 * npts));
  const unsigned g1 = (unsigned)((npts + 255) / 256);
  hipLaunchKernelGGL(k_thc_iota, dim3(g1), dim3(256), 0, C.stream, w.pcand.p, npts);
  hipLaunchKernelGGL(k_thc_fill, dim3(g1), dim3(256), 0, C.stream, w.pD.p, npts, 1.0);
  PivotState st0{};
  PivotState *dst = (PivotState *)w.pstate.p;
  int cur = 0, scur = 0;
  bool converged = false, schur = false;
  long long lbase = 0;   // first vector index held in L
  const double one = 1.0, zero = 0.0, mone = -1.0;
  std::vector<double> hD;
  std::vector<long long> map;
  for (long long it = 0; it < max_iters; ++it) {
    if (!schur) {
      // orbitals of the candidates, normalized (thc_Chol_Pivots :893-895)
      hipLaunchKernelGGL(k_thc_gather_xyz, dim3((unsigned)((ncand + 255) / 256)), dim3(256), 0, C.stream, dxyz,
                         w.pcand.p, ncand, w.pxyz.p);
      w.pX.alloc((size_t)(ncand * nao));
      thc_orbitals(C, w.pxyz.p, ncand, 1, w.pX.p, ncand);
      if ((double)ncand * (double)ncand * 8.0 <= schur_bytes) {
        // Shat = (X X^T)^2 - L L^T over the candidates; L is dropped
        w.pS[scur].alloc((size_t)(ncand * ncand));
        THC_BLAS(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_transpose, ncand, ncand, nao, &one,
                               w.pX.p, ncand, w.pX.p, ncand, &zero, w.pS[scur].p, ncand));
        hipLaunchKernelGGL(k_thc_square, dim3((unsigned)((ncand * ncand + 255) / 256)), dim3(256), 0, C.stream,
                           w.pS[scur].p, ncand, ncand);
        GI_HIP(hipGetLastError());
        if (npiv > 0)
          THC_BLAS(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_transpose, ncand, ncand, npiv,
                                 &mone, w.pL[cur].p, ncand, w.pL[cur].p, ncand, &one, w.pS[scur].p, ncand));
        schur = true;
        lbase = npiv;
      }
    }
    const long long max_j = std::min(block_dim, npts - npiv);
    if (schur)
      w.pL[cur].alloc((size_t)(ncand * std::max<long long>(max_j, 1)));
    else
      w.pL[cur].grow_keep((size_t)(ncand * (npiv + max_j)), C.stream);
    if (schur) lbase = npiv;   // the block's vectors only
    // the state's done flag starts clear for every block
    st0.npiv = npiv;
    st0.done = 0;
    C.h2d(w.pstate.p, &st0, sizeof(st0));
    const unsigned gc = (unsigned)((ncand + 63) / 64);
    const double *Sh = schur ? w.pS[scur].p : nullptr;
    for (long long k = 0; k < max_j; ++k)
      hipLaunchKernelGGL(k_thc_pivot_step, dim3(gc), dim3(1024), 0, C.stream, w.pX.p, ncand, nao, Sh, w.pL[cur].p,
                         lbase, npiv + k, Dbuf[(dcur + k) & 1], Dbuf[(dcur + k + 1) & 1], w.pcand.p, w.ppiv.p, dst,
                         thresh, thresh_red);
    GI_HIP(hipGetLastError());
    C.read(&st0, w.pstate.p, sizeof(st0));
    const long long nblk = st0.npiv - npiv;
    dcur = (int)((dcur + nblk) & 1);   // the buffer the block's last vector wrote
    npiv = st0.npiv;
    nred = st0.nred;
    // Schur form: Shat -= L_blk L_blk^T
    if (schur && nblk > 0)
      THC_BLAS(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_transpose, ncand, ncand, nblk, &mone,
                             w.pL[cur].p, ncand, w.pL[cur].p, ncand, &one, w.pS[scur].p, ncand));
    // compression: keep the candidates with D > thresh, in order
    hD.resize((size_t)ncand);
    C.read(hD.data(), Dbuf[dcur], sizeof(double) * ncand);
    map.clear();
    for (long long g = 0; g < ncand; ++g)
      if (hD[g] > thresh) map.push_back(g);
    if (map.empty()) {
      converged = true;
      break;
    }
    const long long nnew = (long long)map.size();
    w.pmap.upload(map, C.stream);
    const int nxt = 1 - cur;
    const long long nkeep = schur ? 0 : npiv;   // vectors carried to the next block
    w.pL[nxt].alloc((size_t)(nnew * std::max<long long>(nkeep, 1)));
    DBuf<long long> cn;
    cn.alloc((size_t)nnew);
    hipLaunchKernelGGL(k_thc_compress, dim3((unsigned)((nnew + 255) / 256), ygrid(nkeep)), dim3(256), 0, C.stream,
                       w.pmap.p, nnew, ncand, nkeep, w.pL[cur].p, w.pL[nxt].p, Dbuf[dcur], Dbuf[1 - dcur], w.pcand.p,
                       cn.p);
    dcur = 1 - dcur;
    GI_HIP(hipGetLastError());
    if (schur) {
      w.pS[1 - scur].alloc((size_t)(nnew * nnew));
      hipLaunchKernelGGL(k_thc_gather_square, dim3((unsigned)((nnew + 255) / 256), ygrid(nnew)), dim3(256), 0,
                         C.stream, w.pmap.p, nnew, ncand, w.pS[scur].p, w.pS[1 - scur].p);
      GI_HIP(hipGetLastError());
      scur = 1 - scur;
    }
    GI_HIP(hipMemcpyAsync(w.pcand.p, cn.p, sizeof(long long) * nnew, hipMemcpyDeviceToDevice, C.stream));
    cur = nxt;
    ncand = nnew;
  }
  if (!converged) throw Error{85, "Cholesky decomposistion failed to converge"};
  if (!(qr_thresh_reduced > 0.0)) nred = npiv;
  pivots.resize((size_t)npiv);
  if (npiv > 0) C.read(pivots.data(), w.ppiv.p, sizeof(long long) * npiv);
}

}  // namespace gi
