// AO -> MO transform of the Cholesky vectors on the device,
//   Rkab(k, a, b) = sum_{p,q} L^k_{pq} CA(p, a) CB(q, b),
// the operation of chol_gammcor_MOTransf_2 (Cholesky_Gammcor.f90:143-332),
// recast from the reference's rank-1 updates (dger) and matrix-vector
// products (dgemv) into two FP64 GEMMs per vector block:
//   1. unpack R(k, pq) of a block of vectors into dense symmetric L^k
//      (pairs screened out of BASE are zero, as in the reference's Step 1
//      loop which visits BASE pairs only),
//   2. W(k, a, p) = sum_q L^k(q, p) CA(q, a)   strided-batched DGEMM over p,
//   3. Rkab(k a, b) = sum_p W(k a, p) CB(p, b)  one DGEMM.
#include <algorithm>
#include <cmath>
#include "gi_context.h"

namespace gi {

#define RB_CHECK(x)                                                                               \
  do {                                                                                           \
    rocblas_status s_ = (x);                                                                     \
    if (s_ != rocblas_status_success)                                                            \
      throw Error{2, std::string("rocBLAS error ") + rocblas_status_to_string(s_) + " at " + __FILE__ + ":" + \
                         std::to_string(__LINE__)};                                              \
  } while (0)

// L[kk + kb (q + nao p)] = Rkpq[(k0 + kk) + nvec * map[q + nao p]] or 0
__global__ void k_unpack(double *__restrict__ L, const double *__restrict__ R, const long long *__restrict__ map,
                         long long nvec, long long k0, long long kb, long long nao) {
  long long kk = threadIdx.x + (long long)blockIdx.x * blockDim.x;
  long long qp = blockIdx.y + (long long)gridDim.y * blockIdx.z;
  if (kk >= kb || qp >= nao * nao) return;
  long long idx = map[qp];
  L[kk + kb * qp] = idx >= 0 ? R[(k0 + kk) + nvec * idx] : 0.0;
}

// zero Rkab(:, a, :) for the a columns in `cols`
__global__ void k_zero_acols(double *Rkab, long long nvec, long long na, long long nb, const long long *cols,
                             long long ncols) {
  long long k = threadIdx.x + (long long)blockIdx.x * blockDim.x;
  long long c = blockIdx.y, b = blockIdx.z;
  if (k >= nvec || c >= ncols || b >= nb) return;
  Rkab[k + nvec * (cols[c] + na * b)] = 0.0;
}

void mo_transform(Context &C, const double *dCA, long long lda, long long a0, long long a1, const double *dCB,
                  long long ldb, long long b0, long long b1, long long max_buffer_mb, long long flags,
                  double *dRkab) {
  CholState &S = C.chol;
  if (!S.valid || !S.rkpq.p) throw Error{50, "Cholesky vectors are not available (run gi_chol_build first)"};
  const long long nvec = S.nvecs, nao = C.nao, msd = S.max_subset_dim;
  const long long NA = a1 - a0 + 1, NB = b1 - b0 + 1;
  if (NA <= 0 || NB <= 0 || a0 < 1 || b0 < 1) throw Error{51, "Invalid dimension of Rkab"};
  if (lda < nao || ldb < nao) throw Error{52, "Invalid number of atomic orbitals in CA/CB"};
  // pass structure of the reference (Cholesky_Gammcor.f90:171-185, 190-197)
  long long maxbuf = (max_buffer_mb * 1024 * 1024) / 8;
  long long maxnk = std::min(NA, maxbuf / std::max<long long>(nvec * nao, 1));
  if (!(maxnk > 0)) throw Error{53, "Buffer size too small for Rpq->Rab"};
  std::vector<char> done(NA, 0);
  long long npasses = NA / maxnk + (NA % maxnk > 0);
  for (long long k = 1; k <= npasses; ++k) {
    long long a0k = (k - 1) * maxnk + a0, a1k = std::min(k * maxnk, a1);
    for (long long a = a0k; a <= a1k; ++a) done[a - a0] = 1;
  }
  // AO pair -> Rkpq column map (cached per Cholesky build)
  if (!S.mo_map.p) {
    std::vector<long long> map(nao * nao, -1);
    for (long long s = 0; s < S.nshell_pairs; ++s) {
      int A = (int)S.shell_pairs[2 * s] - 1, B = (int)S.shell_pairs[2 * s + 1] - 1;
      long long X = S.shell_pair_loc[3 * s + 2], col = S.shell_pair_loc[3 * s + 1] - 1;
      int na = C.shell_nang[A], nb = C.shell_nang[B];
      long long base = (X - 1) * msd + col;
      long long q = 0;
      for (int b = 0; b < nb; ++b)
        for (int a = (A == B ? b : 0); a < na; ++a) {
          long long p1 = C.shell_loc[A] + a, p2 = C.shell_loc[B] + b;
          map[p1 + nao * p2] = base + q;
          map[p2 + nao * p1] = base + q;
          ++q;
        }
    }
    S.mo_map.upload(map, C.stream);
  }
  // vector blocks: keep the dense L^k block under ~4 GiB
  long long kb = std::max<long long>(1, (long long)((4.0 * (1ull << 30)) / (8.0 * (double)nao * nao)));
  kb = std::min(kb, nvec);
  if (kb > 64) kb -= kb % 64;
  DBuf<double> dL, dW;
  dL.alloc((size_t)kb * nao * nao);
  dW.alloc((size_t)nvec * NA * nao);
  const double one = 1.0, zero = 0.0;
  for (long long k0 = 0; k0 < nvec; k0 += kb) {
    long long kbl = std::min(kb, nvec - k0);
    long long nqp = nao * nao;
    dim3 grid((unsigned)((kbl + 63) / 64), (unsigned)std::min<long long>(nqp, 65535),
              (unsigned)((nqp + 65534) / 65535));
    hipLaunchKernelGGL(k_unpack, grid, dim3(64), 0, C.stream, dL.p, S.rkpq.p, S.mo_map.p, nvec, k0, kbl, nao);
    RB_CHECK(rocblas_dgemm_strided_batched(C.blas, rocblas_operation_none, rocblas_operation_none, kbl, NA, nao,
                                           &one, dL.p, kbl, (rocblas_stride)(kbl * nao), dCA + (a0 - 1) * lda, lda, 0,
                                           &zero, dW.p + k0, nvec, (rocblas_stride)(nvec * NA), nao));
  }
  RB_CHECK(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_none, nvec * NA, NB, nao, &one, dW.p,
                         nvec * NA, dCB + (b0 - 1) * ldb, ldb, &zero, dRkab, nvec * NA));
  if (!(flags & 1)) {
    std::vector<long long> skip;
    for (long long a = 0; a < NA; ++a)
      if (!done[a]) skip.push_back(a);
    if (!skip.empty()) {
      DBuf<long long> ds;
      ds.upload(skip, C.stream);
      dim3 g((unsigned)((nvec + 255) / 256), (unsigned)skip.size(), (unsigned)NB);
      hipLaunchKernelGGL(k_zero_acols, g, dim3(256), 0, C.stream, dRkab, nvec, NA, NB, ds.p, (long long)skip.size());
      GI_HIP(hipStreamSynchronize(C.stream));
    }
  }
  GI_HIP(hipStreamSynchronize(C.stream));
}

}  // namespace gi
