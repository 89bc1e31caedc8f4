// AO -> MO transform of the Cholesky vectors on the device,
//   Rkab(k, a, b) = sum_{p,q} L^k_{pq} CA(p, a) CB(q, b),
// the operation of chol_gammcor_MOTransf_2 (Cholesky_Gammcor.f90:143-332).
// The reference's first step adds R_k(pq) CA(p,a) into W(:,a,q) and
// R_k(pq) CA(q,a) into W(:,a,p) pair by pair (two dger per orbital pair,
// :215-292).  Here the same sum runs pair-list by pair-list: for every AO q,
// the orbital pairs that contain q (p, column of Rkpq) form a gathered
// product
//   W(k, a, q) = sum_{j in pairs(q)} R(k, col_j) CA(p_j, a)
// on FP64 MFMA (k_mo_half), reading Rkpq in place (no dense L^k), then
//   Rkab(k a, b) = sum_q W(k a, q) CB(q, b)
// is one DGEMM (the reference's second step, dgemv per b, :295-332).
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

typedef double mo_d4 __attribute__((ext_vector_type(4)));

// W(k, a, q) for one AO q (blockIdx.x), 64 vectors k0.. (blockIdx.y) and 64
// MO columns a0.. (blockIdx.z).  MFMA f64 16x16x4: A = CA^T rows of the
// partners p (a x j), B = the Rkpq columns of the pairs (j x k), so the
// result tiles are (a, k) with k fastest across lanes (coalesced stores).
// CT(a, p) = CA(p, a0 + a) with leading dimension ldct (>= NA rounded up to 64).
__global__ void __launch_bounds__(64) k_mo_half(const double *__restrict__ R, long long nvec,
                                                const long long *__restrict__ adj_off,
                                                const long long *__restrict__ adj_col,
                                                const int *__restrict__ adj_p, const double *__restrict__ CT,
                                                long long ldct, long long NA, double *__restrict__ W) {
  const int lane = threadIdx.x, i16 = lane & 15, g = lane >> 4;
  const long long q = blockIdx.x;
  const long long k0 = (long long)blockIdx.y * 64, a0 = (long long)blockIdx.z * 64;
  const long long j0 = adj_off[q], j1 = adj_off[q + 1];
  mo_d4 acc[4][4];
#pragma unroll
  for (int at = 0; at < 4; ++at)
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) acc[at][kt] = mo_d4{0.0, 0.0, 0.0, 0.0};
  for (long long j = j0; j < j1; j += 4) {
    const long long jj = j + g;           // this lane's partner in the k-step
    const bool ok = jj < j1;
    const long long col = ok ? adj_col[jj] : 0;
    const int p = ok ? adj_p[jj] : 0;
    double av[4], bv[4];
#pragma unroll
    for (int at = 0; at < 4; ++at) {
      const long long a = a0 + 16 * at + i16;
      av[at] = (ok && a < NA) ? CT[a + ldct * p] : 0.0;
    }
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      const long long k = k0 + 16 * kt + i16;
      bv[kt] = (ok && k < nvec) ? R[k + nvec * col] : 0.0;
    }
#pragma unroll
    for (int at = 0; at < 4; ++at)
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) acc[at][kt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[at], bv[kt], acc[at][kt], 0, 0, 0);
  }
  // element (i, j) of a result tile sits in lane j + 16 (i % 4), register i / 4
#pragma unroll
  for (int at = 0; at < 4; ++at)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const long long a = a0 + 16 * at + 4 * r + g;
      if (a >= NA) continue;
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) {
        const long long k = k0 + 16 * kt + i16;
        if (k < nvec) W[k + nvec * (a + NA * q)] = acc[at][kt][r];
      }
    }
}

// CT(a, p) = CA(p, a0 + a), zero-padded to ldct rows
__global__ void k_mo_ct(double *__restrict__ CT, long long ldct, const double *__restrict__ CA, long long lda,
                        long long a0, long long NA, long long nao) {
  const long long a = threadIdx.x + (long long)blockIdx.x * blockDim.x;
  const long long p = blockIdx.y;
  if (a < ldct && p < nao) CT[a + ldct * p] = a < NA ? CA[p + lda * (a0 - 1 + a)] : 0.0;
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
  const long long npasses = NA / maxnk + (NA % maxnk > 0);
  for (long long k = 1; k <= npasses; ++k) {
    long long a0k = (k - 1) * maxnk + a0, a1k = std::min(k * maxnk, a1);
    for (long long a = a0k; a <= a1k; ++a) done[a - a0] = 1;
  }
  // pair lists per AO (cached per Cholesky build): for AO q every BASE
  // orbital pair (p, q) or (q, p) with its Rkpq column (X - 1) msd + position
  if (!S.mo_adj_off.p) {
    std::vector<std::vector<std::pair<int, long long>>> adj(nao);
    for (long long s = 0; s < S.nshell_pairs; ++s) {
      const int A = (int)S.shell_pairs[2 * s] - 1, B = (int)S.shell_pairs[2 * s + 1] - 1;
      const long long X = S.shell_pair_loc[3 * s + 2], pos0 = S.shell_pair_loc[3 * s + 1] - 1;
      const int na = C.shell_nang[A], nb = C.shell_nang[B];
      long long qq = 0;
      for (int b = 0; b < nb; ++b)
        for (int a = (A == B ? b : 0); a < na; ++a) {
          const int p1 = C.shell_loc[A] + a, p2 = C.shell_loc[B] + b;
          const long long col = (X - 1) * msd + pos0 + qq;
          adj[p2].push_back({p1, col});
          if (p1 != p2) adj[p1].push_back({p2, col});
          ++qq;
        }
    }
    std::vector<long long> off(nao + 1, 0), cols;
    std::vector<int> ps;
    for (long long q = 0; q < nao; ++q) {
      off[q + 1] = off[q] + (long long)adj[q].size();
      for (auto &e : adj[q]) {
        ps.push_back(e.first);
        cols.push_back(e.second);
      }
    }
    S.mo_adj_off.upload(off, C.stream);
    S.mo_adj_col.upload(cols, C.stream);
    S.mo_adj_p.upload(ps, C.stream);
  }
  // passes over a-column ranges, W(k, a, q) of at most maxnk columns
  // (the reference's NCholesky*MaxNk*NAO buffer, :186); GI_MO_ALL_COLUMNS
  // takes contiguous ranges covering a0..a1, otherwise the reference's pass
  // bounds (columns they skip stay zero)
  std::vector<std::pair<long long, long long>> passes;
  for (long long k = 1; k <= npasses; ++k) {
    long long a0k, a1k;
    if (flags & 1) {
      a0k = a0 + (k - 1) * maxnk;
      a1k = std::min(a0k + maxnk - 1, a1);
    } else {
      a0k = (k - 1) * maxnk + a0;
      a1k = std::min(k * maxnk, a1);
    }
    if (a1k >= a0k) passes.push_back({a0k, a1k});
  }
  long long nk_max = 0;
  for (auto &pr : passes) nk_max = std::max(nk_max, pr.second - pr.first + 1);
  const long long ldct = (std::max<long long>(nk_max, 1) + 63) / 64 * 64;
  DBuf<double> dCT, dW;
  dCT.alloc((size_t)ldct * nao);
  dW.alloc((size_t)nvec * std::max<long long>(nk_max, 1) * nao);
  const double one = 1.0, zero = 0.0;
  for (auto &pr : passes) {
    const long long a0k = pr.first, Nk = pr.second - pr.first + 1;
    hipLaunchKernelGGL(k_mo_ct, dim3((unsigned)(ldct / 64), (unsigned)nao), dim3(64), 0, C.stream, dCT.p, ldct,
                       dCA, lda, a0k, Nk, nao);
    hipLaunchKernelGGL(k_mo_half, dim3((unsigned)nao, (unsigned)((nvec + 63) / 64), (unsigned)((Nk + 63) / 64)),
                       dim3(64), 0, C.stream, S.rkpq.p, nvec, S.mo_adj_off.p, S.mo_adj_col.p, S.mo_adj_p.p, dCT.p,
                       ldct, Nk, dW.p);
    GI_HIP(hipGetLastError());
    // Rkab(k, a, b), a in this pass: rows nvec (a0k - a0) .. of each b column
    double *dst = dRkab + nvec * (a0k - a0);
    if (flags & 2)
      split_bf16_gemm(C, rocblas_operation_none, rocblas_operation_none, nvec * Nk, NB, nao, 1.0, dW.p, nvec * Nk,
                      dCB + (b0 - 1) * ldb, ldb, 0.0, dst, nvec * NA);
    else
      RB_CHECK(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_none, nvec * Nk, NB, nao, &one, dW.p,
                             nvec * Nk, dCB + (b0 - 1) * ldb, ldb, &zero, dst, nvec * NA));
  }
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
