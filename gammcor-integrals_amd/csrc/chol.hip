// Two-step Cholesky decomposition of the Coulomb matrix on one MI355X
// (Folkestad, Kjonstad, Koch, JCP 150, 194112 (2019)), following the
// reference driver step by step:
//   chol_gammcor_CoulombMatrix       Cholesky_Gammcor.f90:380-441
//   chol2_Step1_FullInterface        TwoStepCholesky.f90:74-167
//   chol2_Vdiag/DefineBase/Subsets   TwoStepCholesky_Step1.f90:202-530
//   chol2_Pivots (macro/micro loop)  TwoStepCholesky_Step1.f90:884-1103
//   chol2_Step2                      TwoStepCholesky.f90:243-364
//   chol2_PivotShellPairs            TwoStepCholesky_Step2.f90:422-534
//   chol2_FullDimVectors_Batch       TwoStepCholesky.f90:378-416
//
// Everything O(N^3..N^4) runs on the device: ERIs (class kernels), the
// candidate compaction, M - L^T L (rocBLAS DGEMM), the micro-iteration
// column update + diagonal update + batch argmax (own kernels), the Cholesky
// factor of V(P,P) and its inverse (own recursive factorization: LDS leaf
// kernel + rocBLAS DGEMMs) and Inv_L V^T (triangular-blocked DGEMMs).  The host keeps the
// bookkeeping whose order fixes the pivot sequence: the shell-pair sorts
// (Singleton quicksort, identical permutation for tied keys), the BASE
// screening, the batch selection and the strict-> first-maximum rule.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include "gi_context.h"

namespace gi {

void dsort0(double *dx, long long *dy, long long n, int kflag);   // sort.cpp

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

#define ROCBLAS_CHECK(x)                                                         \
  do {                                                                          \
    rocblas_status s_ = (x);                                                    \
    if (s_ != rocblas_status_success)                                           \
      throw Error{2, std::string("rocBLAS/rocSOLVER error ") + rocblas_status_to_string(s_) + \
                         " at " + __FILE__ + ":" + std::to_string(__LINE__)};   \
  } while (0)

// ------------------------------------------------------------------ kernels
// out[c] = in[src[c]] for a column vector; for a matrix with nrows rows:
// out[r + ldo c] = in[r + ldi src[c]] (gather of whole columns)
__global__ void k_gather_rows_of_cols(double *__restrict__ out, long long ldo, const double *__restrict__ in,
                                      long long ldi, const long long *__restrict__ src, long long ncols,
                                      long long nrows) {
  // out is (ncols_new x nvec): out[c + ldo*r] = in[src[c] + ldi*r]
  long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long r = blockIdx.y;
  for (; r < nrows; r += gridDim.y)
    if (c < ncols) out[c + ldo * r] = in[src[c] + ldi * r];
}

__global__ void k_segmax(const double *__restrict__ D, const long long *__restrict__ loc,
                         const int *__restrict__ dim, long long n, double *__restrict__ out) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double *p = D + loc[k];
  double m = p[0];
  for (int i = 1; i < dim[k]; ++i) m = p[i] > m ? p[i] : m;
  out[k] = m;
}

// Near-tie canonicalization of comparison keys (same rule as the oracle):
// values equal in exact arithmetic differ in their last bits with a sign set
// by each implementation's rounding; comparing quantized keys makes the
// pivot order a function of the mathematics.  The quanta sit well above the
// disagreement of two correct ERI evaluations (Boys-function and summation
// order differences reach 1e-11 relative on small diagonals, 1e-13 absolute
// on residuals) and well below tau.  BASE sort keys: mantissa rounded to
// KEY_REL_BITS bits.  Residual diagonals: absolute quantum
// delta = min(2^(floor(log2 MaxDrs) - KEY_ABS_BITS), 2^(floor(log2 tau) -
// KEY_TAU_BITS)), a power of two so x/delta is exact.  Only orderings use the
// quantized keys; the thresholds (tau for candidates and convergence, Dmin
// for the batch and the micro iterations) compare raw values, as the
// reference does.
constexpr int KEY_REL_BITS = 26;
constexpr int KEY_ABS_BITS = 32;
constexpr int KEY_TAU_BITS = 8;
static double qrel(double x) {
  if (x == 0.0 || !std::isfinite(x)) return x;
  int e;
  double m = std::frexp(x, &e);
  return std::ldexp(std::rint(std::ldexp(m, KEY_REL_BITS + 1)), e - KEY_REL_BITS - 1);
}
__host__ __device__ static inline double qabs(double x, double inv_delta) { return rint(x * inv_delta); }

// ---- micro iterations restricted to the batch rows -------------------------
// The pivot of every micro iteration is chosen among the batch elements, and
// its row of L is a batch row, so the sequential part of chol2_Pivots only
// needs L and D on the batch rows (at most 1000 of them: the batch limit of
// chol2_DefineBatch).  The loop runs on the device in panels of MICRO_PANEL
// steps: one workgroup of 1024 lanes (one lane per batch row) runs a panel
// with the panel's columns in LDS, computing each column from the batch block
// A, which holds M(batch, batch) minus the contributions of all earlier
// panels; k_micro_update then subtracts the panel's rank-MICRO_PANEL term
// from A.  The first-maximum search stays inside the workgroup, so a panel is
// one launch with no host round trip.  The remaining candidate rows are
// filled afterwards in one triangular solve (see chol_build).
constexpr int MICRO_PANEL = 16;
constexpr int MICRO_ROWS = 1024;

struct MicroState {
  int done;       // loop has stopped
  int j;          // columns computed in this macro iteration
  int e;          // batch index of the current pivot
  int p0;         // first column of the current panel
  double val;     // raw residual diagonal of the current pivot
};

// Device work arrays of gi_chol_build, kept in the context across builds and
// grown only: a repeated build (same or smaller system) neither allocates nor
// frees device memory.  Stream-ordered frees of multi-GB blocks between the
// macro iterations stalled the host for up to 0.7 s in long runs.
struct CholWork {
  DBuf<double> D, Dn, L, Ln, M, LB, tmp, bm, Mb, Lb, Db, LP;   // Step 1
  DBuf<int> piv;
  DBuf<MicroState> state;
  DBuf<long long> src, loc, bd;
  DBuf<unsigned char> excl;
  DBuf<int> dim;
  DBuf<int> info;            // factorization of V(P, P)
  DBuf<double> X, W;
  DBuf<double> Wv;           // Vabrs staging of the full-dimension vectors
};
static CholWork &chol_work(Context &C) {
  if (!C.cw) C.cw = new CholWork();
  return *C.cw;
}
void release_chol_work(Context &C) {
  delete C.cw;
  C.cw = nullptr;
}


__device__ __forceinline__ void argmax_merge(double &k1, double &v1, long long &i1, double k2, double v2,
                                             long long i2) {
  bool take = (i2 >= 0) && (i1 < 0 || k2 > k1 || (k2 == k1 && i2 < i1));
  if (take) { k1 = k2; v1 = v2; i1 = i2; }
}

template <int BS>
__device__ void block_argmax(double &k, double &v, long long &i) {
  __shared__ double sk[BS], sv[BS];
  __shared__ long long si[BS];
  sk[threadIdx.x] = k;
  sv[threadIdx.x] = v;
  si[threadIdx.x] = i;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      double kk = sk[threadIdx.x], vv = sv[threadIdx.x];
      long long ii = si[threadIdx.x];
      argmax_merge(kk, vv, ii, sk[threadIdx.x + s], sv[threadIdx.x + s], si[threadIdx.x + s]);
      sk[threadIdx.x] = kk; sv[threadIdx.x] = vv; si[threadIdx.x] = ii;
    }
    __syncthreads();
  }
  k = sk[0]; v = sv[0]; i = si[0];
}

// gather the batch block Mb(e, c) = M(bd(e), c), the batch diagonals, and the
// initial first maximum
// With Mb (precision form) the batch diagonals are taken from the updated
// block itself, so that a pivot's residual vanishes within the recurrence, and
// rows chosen in earlier macro iterations (excl) start at zero: the split
// trailing update leaves them O(2^-24 |V|) instead of exactly zero.
__global__ void __launch_bounds__(1024) k_micro_init(const double *__restrict__ D, const long long *__restrict__ bd,
                                                      long long nb, double inv_delta, double *__restrict__ Db,
                                                      MicroState *st, int nsteps, const double *__restrict__ Mb,
                                                      const unsigned char *__restrict__ excl) {
  double bk = inv_delta > 0.0 ? qabs(-1.0, inv_delta) : -1.0, bv = -1.0;
  long long bi = -1;
  for (long long e = threadIdx.x; e < nb; e += blockDim.x) {
    double v = Mb ? (excl[e] ? 0.0 : Mb[e + nb * e]) : D[bd[e]];
    Db[e] = v;
    double k = inv_delta > 0.0 ? qabs(v, inv_delta) : v;
    argmax_merge(bk, bv, bi, k, v, e);
  }
  block_argmax<1024>(bk, bv, bi);
  if (threadIdx.x == 0) {
    st->j = 0;
    st->p0 = 0;
    st->e = (int)bi;
    st->val = bv;
    st->done = (bi < 0 || nsteps <= 0) ? 1 : 0;
  }
}

__global__ void k_gather_block(double *__restrict__ Mb, const double *__restrict__ M, long long ldM,
                               const long long *__restrict__ bd, long long nb) {
  long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long c = blockIdx.y;
  if (e < nb && c < nb) Mb[e + nb * c] = M[bd[e] + ldM * c];
}

// first maximum over a 1024-lane workgroup (key descending, index ascending;
// the order of merging does not change the result).  Ends with a barrier.
__device__ void wg_argmax_1024(double &k, double &v, int &i) {
  __shared__ double sk[16], sv[16];
  __shared__ int si[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double k2 = __shfl_xor(k, off), v2 = __shfl_xor(v, off);
    const int i2 = __shfl_xor(i, off);
    const bool take = (i2 >= 0) && (i < 0 || k2 > k || (k2 == k && i2 < i));
    if (take) { k = k2; v = v2; i = i2; }
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) { sk[w] = k; sv[w] = v; si[w] = i; }
  __syncthreads();
  if (w == 0) {
    if (lane < 16) { k = sk[lane]; v = sv[lane]; i = si[lane]; }
    else { i = -1; }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) {
      const double k2 = __shfl_xor(k, off), v2 = __shfl_xor(v, off);
      const int i2 = __shfl_xor(i, off);
      const bool take = (i2 >= 0) && (i < 0 || k2 > k || (k2 == k && i2 < i));
      if (take) { k = k2; v = v2; i = i2; }
    }
    if (lane == 0) { sk[0] = k; sv[0] = v; si[0] = i; }
  }
  __syncthreads();
  k = sk[0]; v = sv[0]; i = si[0];
  __syncthreads();
}

// One panel of micro iterations (chol2_Lj + D update + BatchMax,
// TwoStepCholesky_Step1.f90:1027-1064) on the batch rows, nb <= MICRO_ROWS,
// by a workgroup of MICRO_ROWS lanes (lane i = batch row i):
//   L(i, j) = (A(i, e) - sum_{p0 <= q < j} L(i, q) L(e, q)) / sqrt(D(e)),
//   D(i) -= L(i, j)^2,  e <- first maximum of D (quantized keys).
// (j, e, val) carry the loop state; lp[step][i] receives the panel's columns.
__device__ void micro_panel_steps(double (&lp)[MICRO_PANEL][MICRO_ROWS], const double *__restrict__ A,
                                  double *__restrict__ Lb, double *__restrict__ Db, long long nb, int nsteps,
                                  double dmin, double inv_delta, int *__restrict__ piv, int &j, int &e, double &val,
                                  int &done) {
  const int i = threadIdx.x;
  const bool row = i < nb;
  done = 0;
  for (int step = 0; step < MICRO_PANEL; ++step) {
    const double sq = sqrt(val);
    double bk = inv_delta > 0.0 ? qabs(-1.0, inv_delta) : -1.0, bv = -1.0;
    int bi = -1;
    if (row) {
      double s = A[i + nb * e];
      for (int q = 0; q < step; ++q) s -= lp[q][i] * lp[q][e];
      const double l = s / sq;
      lp[step][i] = l;
      Lb[i + nb * j] = l;
      const double d = Db[i] - l * l;
      Db[i] = d;
      bk = inv_delta > 0.0 ? qabs(d, inv_delta) : d;
      bv = d;
      bi = i;
    }
    wg_argmax_1024(bk, bv, bi);   // its barriers also publish the panel column
    if (threadIdx.x == 0) piv[j] = e;
    ++j;
    e = bi;
    val = bv;
    done = (bi < 0 || bv <= dmin || j >= nb || j >= nsteps) ? 1 : 0;
    if (done) break;
  }
}

__global__ void __launch_bounds__(MICRO_ROWS) k_micro_panel(MicroState *st, const double *__restrict__ A,
                                                            double *__restrict__ Lb, double *__restrict__ Db,
                                                            long long nb, int nsteps, double dmin, double inv_delta,
                                                            int *__restrict__ piv) {
  __shared__ double lp[MICRO_PANEL][MICRO_ROWS];
  __shared__ int s_done;
  if (threadIdx.x == 0) s_done = st->done;
  __syncthreads();
  if (s_done) return;
  int j = st->j, e = st->e, done = 0;
  double val = st->val;
  const int p0 = j;
  micro_panel_steps(lp, A, Lb, Db, nb, nsteps, dmin, inv_delta, piv, j, e, val, done);
  if (threadIdx.x == 0) {
    st->p0 = p0;
    st->j = j;
    st->e = e;
    st->val = val;
    st->done = done;
  }
}

// A(i, c) -= sum_{q < np} lp[q][i] lp[q][c] over the nb x nb block, the panel
// columns read from LDS: 16 x 16 tiles of FP64 16x16x4 MFMA, one wavefront
// per tile (MICRO_ROWS / 64 of them in turn).  The update is symmetric, so
// the instruction's first operand takes the tile's columns c and its second
// the rows i; lane l then holds rows i0 + (l & 15) of columns
// c0 + 4 qq + (l >> 4), sixteen consecutive doubles of A per column.
__device__ void micro_update_lds(const double (&lp)[MICRO_PANEL][MICRO_ROWS], double *__restrict__ A, long long nb,
                                 int np) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, n = l & 15, g = l >> 4;
  const int ntr = (int)((nb + 15) / 16);
  const int kst = (np + 3) / 4;
  for (int t = w; t < ntr * ntr; t += MICRO_ROWS / 64) {
    const int ti = t % ntr, tc = t / ntr;
    const int i = 16 * ti + n, cm = 16 * tc + n;
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int ks = 0; ks < kst; ++ks) {
      const int q = 4 * ks + g;
      const double la = lp[q][cm < nb ? cm : 0], lb = lp[q][i < nb ? i : 0];
      const double a = (q < np && cm < nb) ? la : 0.0;
      const double b = (q < np && i < nb) ? lb : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const int c = 16 * tc + 4 * qq + g;
      if (i < nb && c < nb) A[i + nb * c] -= acc[qq];
    }
  }
}

// All micro iterations of one macro iteration in one launch: panels of
// MICRO_PANEL steps, each followed by its rank update of the batch block from
// LDS, until the loop stops (at most ceil(nsteps / MICRO_PANEL) panels).
// For batch blocks up to a size where one workgroup's update is cheaper than
// two launches per panel (and a state read-back every few panels).
__global__ void __launch_bounds__(MICRO_ROWS) k_micro_all(MicroState *st, double *__restrict__ A,
                                                          double *__restrict__ Lb, double *__restrict__ Db,
                                                          long long nb, int nsteps, double dmin, double inv_delta,
                                                          int *__restrict__ piv) {
  __shared__ double lp[MICRO_PANEL][MICRO_ROWS];
  __shared__ int s_done;
  if (threadIdx.x == 0) s_done = st->done;
  __syncthreads();
  if (s_done) return;
  int j = st->j, e = st->e, done = 0, p0 = j;
  double val = st->val;
  const int npanel = (nsteps + MICRO_PANEL - 1) / MICRO_PANEL;
  for (int pnl = 0; pnl < npanel; ++pnl) {
    p0 = j;
    micro_panel_steps(lp, A, Lb, Db, nb, nsteps, dmin, inv_delta, piv, j, e, val, done);
    if (done) break;
    micro_update_lds(lp, A, nb, j - p0);
    __syncthreads();   // A updated, lp free for the next panel
  }
  if (threadIdx.x == 0) {
    st->p0 = p0;
    st->j = j;
    st->e = e;
    st->val = val;
    st->done = 1;
  }
}

// A(i, c) -= sum_{q = p0}^{j-1} L(i, q) L(c, q) for the panel just finished
// (skipped once the loop has stopped: A is not read again).
__global__ void k_micro_update(const MicroState *st, double *__restrict__ A, const double *__restrict__ Lb,
                               long long nb) {
  if (st->done) return;
  const int p0 = st->p0, j = st->j;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long c = blockIdx.y;
  if (i >= nb || c >= nb) return;
  double s = A[i + nb * c];
  for (int q = p0; q < j; ++q) s -= Lb[i + nb * q] * Lb[c + nb * q];
  A[i + nb * c] = s;
}

// D(bd(piv(q))) = 0 for the pivots of this macro iteration (precision form)
__global__ void k_zero_pivot_d(double *__restrict__ D, const long long *__restrict__ bd, const int *__restrict__ piv,
                               long long J) {
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (q < J) D[bd[piv[q]]] = 0.0;
}

// Lt(:, j0+q) = M(:, piv(q))  (right-hand sides of the triangular solve)
__global__ void k_gather_cols(double *__restrict__ Lt, long long ldL, const double *__restrict__ M, long long ldM,
                              const int *__restrict__ piv, long long nrows) {
  long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long q = blockIdx.y;
  if (r < nrows) Lt[r + ldL * q] = M[r + ldM * piv[q]];
}

// L_P(q, r) = Lb(piv(q), r) for r <= q, zero above the diagonal
__global__ void k_build_lp(double *__restrict__ LP, long long J, const double *__restrict__ Lb, long long nb,
                           const int *__restrict__ piv) {
  long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long r = blockIdx.y;
  if (q < J && r < J) LP[q + J * r] = (r <= q) ? Lb[piv[q] + nb * r] : 0.0;
}

// batch rows keep the values of the sequential recurrence; then
// D(i) -= L(i, j0+q)^2 for q = 0..J-1 in order
__global__ void k_finish_rows(double *__restrict__ Lt, long long ldL, long long j0, long long J,
                              const double *__restrict__ Lb, long long nb, const long long *__restrict__ bd,
                              double *__restrict__ D, long long nrows) {
  long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e < nb)
    for (long long q = 0; q < J; ++q) Lt[bd[e] + ldL * (j0 + q)] = Lb[e + nb * q];
}

__global__ void k_update_d(const double *__restrict__ Lt, long long ldL, long long j0, long long J,
                           double *__restrict__ D, long long nrows) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  double d = D[i];
  for (long long q = 0; q < J; ++q) {
    const double l = Lt[i + ldL * (j0 + q)];
    d = d - l * l;
  }
  D[i] = d;
}

__global__ void k_blockmax(const double *__restrict__ D, long long n, double *__restrict__ part) {
  __shared__ double sv[256];
  double m = -INFINITY;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    m = D[i] > m ? D[i] : m;
  sv[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sv[threadIdx.x] = fmax(sv[threadIdx.x], sv[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sv[0];
}


// ---- V = L L^T and L^-1 of the pivot block (real_Cholesky, real_linalg.f90:39-57,
// and real_LowerTriangularInverse, :60-72, called at TwoStepCholesky.f90:344-350).
// LAPACK dpotrf/dtrtri semantics: lower triangle, info = the order of the
// first leading minor that is not positive definite.  Recursive: a diagonal
// block of at most POTF_NB rows is factored and inverted in LDS by one
// workgroup (k_chol_inv_leaf); above that, with the block split in halves,
//   L21 = A21 X11^T,   A22 -= L21 L21^T,   (L22, X22) of A22,   X21 = -X22 (L21 X11),
// X = L^-1, four dgemms per split.  Only X is kept (it is Inv_L).
constexpr int POTF_NB = 64;

__global__ void k_set_int(int *p, int v) { *p = v; }

// rows g msd + dim[g] .. (g + 1) msd - 1 of every column of W (ldw), for the
// ng subsets of a Vabrs group: the padding below each subset's dimension,
// which no ERI block writes
__global__ void k_zero_pad_rows(double *__restrict__ W, long long ldw, long long msd,
                                const long long *__restrict__ dim, long long ncol) {
  const long long g = blockIdx.y;
  const long long r0 = g * msd + dim[g], r1 = (g + 1) * msd;
  for (long long c = blockIdx.x; c < ncol; c += gridDim.x)
    for (long long r = r0 + threadIdx.x; r < r1; r += blockDim.x) W[r + ldw * c] = 0.0;
}

// Cholesky factor L and its inverse X of the diagonal block at k0 (b <= POTF_NB
// rows, lower triangle of V read); writes L (zero above the diagonal) over the
// block of V and X into the block of X.  Does nothing once an earlier block failed.
__global__ void __launch_bounds__(256) k_chol_inv_leaf(double *__restrict__ V, double *__restrict__ X, long long n,
                                                       long long k0, int b, int *__restrict__ info) {
  __shared__ double s[POTF_NB][POTF_NB + 1];
  __shared__ double x[POTF_NB][POTF_NB + 1];
  __shared__ double t[POTF_NB * POTF_NB / 2];   // B A^-1 of the merge step
  __shared__ int s_info;
  const int tid = threadIdx.x;
  if (tid == 0) s_info = *info;
  __syncthreads();
  if (s_info != 0) return;
  // factorization in registers: thread (tr, tc) of a 16 x 16 grid holds the
  // 4 x 4 tile of rows 4 tr .. 4 tr + 3 and columns 4 tc .. 4 tc + 3 (tiles
  // above the diagonal idle); per column j the diagonal goes through LDS, the
  // owners of column j scale it and publish it, every tile takes its rank-1
  // update: two barriers per column.  The block is padded with an identity.
  const int tr = tid & 15, tc = tid >> 4;
  double a[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const int r = 4 * tr + i, c = 4 * tc + jj;
      a[i][jj] = (r < b && c < b && r >= c) ? V[(k0 + r) + n * (k0 + c)] : (r == c ? 1.0 : 0.0);
    }
  for (int e = tid; e < POTF_NB * POTF_NB; e += 256) x[e % POTF_NB][e / POTF_NB] = 0.0;
  __shared__ double s_d;
  __shared__ double s_col[POTF_NB];
  int jend = b;
  for (int j = 0; j < b; ++j) {
    const int tj = j >> 2;
    if (tr == tj && tc == tj) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if ((j & 3) == i) s_d = a[i][i];
    }
    __syncthreads();
    const double d = s_d;
    if (!(d > 0.0)) {   // the same LDS value in every lane: a uniform exit
      jend = j;
      break;
    }
    const double sq = sqrt(d);
    if (tc == tj && tr >= tj) {
#pragma unroll
      for (int jj = 0; jj < 4; ++jj)
        if ((j & 3) == jj) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int r = 4 * tr + i;
            if (r == j) a[i][jj] = sq;
            else if (r > j) a[i][jj] = a[i][jj] / sq;
            if (r >= j) s_col[r] = a[i][jj];
          }
        }
    }
    __syncthreads();
    if (tr >= tc && 4 * tr + 3 > j) {
      double cr[4], cc[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        cr[i] = s_col[4 * tr + i];
        cc[i] = s_col[4 * tc + i];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const int r = 4 * tr + i, c = 4 * tc + jj;
          if (c > j && r >= c) a[i][jj] -= cr[i] * cc[jj];
        }
    }
  }
  if (jend < b) {
    if (tid == 0) *info = (int)(k0 + jend + 1);
    return;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) s[4 * tr + i][4 * tc + jj] = a[i][jj];
  __syncthreads();
  // X = L^-1 (L padded to POTF_NB rows with an identity): the eight 8 x 8
  // diagonal blocks by forward substitution, one column per lane, then pairs
  // of blocks merged, [A 0; B C]^-1 = [A^-1 0; -C^-1 (B A^-1) C^-1], for
  // h = 8, 16, 32, each product spread over all lanes (short dependency chains)
  if (tid < POTF_NB) {
    const int r0 = (tid / 8) * 8, c = tid % 8;
    x[r0 + c][r0 + c] = 1.0 / s[r0 + c][r0 + c];
    for (int r = c + 1; r < 8; ++r) {
      double acc = 0.0;
      for (int k = c; k < r; ++k) acc += s[r0 + r][r0 + k] * x[r0 + k][r0 + c];
      x[r0 + r][r0 + c] = -acc / s[r0 + r][r0 + r];
    }
  }
  __syncthreads();
  for (int h = 8; h < POTF_NB; h *= 2) {
    const int npair = POTF_NB / (2 * h);
    // T = B A^-1 for every pair (A^-1 lower triangular)
    for (int e = tid; e < npair * h * h; e += 256) {
      const int pr = e / (h * h), ij = e % (h * h), i = ij % h, j = ij / h;
      const int a0 = 2 * h * pr, c0 = a0 + h;
      double acc = 0.0;
      for (int k = j; k < h; ++k) acc += s[c0 + i][a0 + k] * x[a0 + k][a0 + j];
      t[e] = acc;
    }
    __syncthreads();
    // X21 = -C^-1 T (C^-1 lower triangular)
    for (int e = tid; e < npair * h * h; e += 256) {
      const int pr = e / (h * h), ij = e % (h * h), i = ij % h, j = ij / h;
      const int a0 = 2 * h * pr, c0 = a0 + h;
      double acc = 0.0;
      for (int k = 0; k <= i; ++k) acc += x[c0 + i][c0 + k] * t[pr * h * h + k + h * j];
      x[c0 + i][a0 + j] = -acc;
    }
    __syncthreads();
  }
  for (int e = tid; e < b * b; e += 256) {
    const int r = e % b, c = e / b;
    V[(k0 + r) + n * (k0 + c)] = r >= c ? s[r][c] : 0.0;
    X[(k0 + r) + n * (k0 + c)] = r >= c ? x[r][c] : 0.0;
  }
}

// ------------------------------------------------------------------ helpers
static int pair_dim_of(const Context &C, int sa, int sb) {
  int na = C.shell_nang[sa], nb = C.shell_nang[sb];
  return sa != sb ? na * nb : na * (na + 1) / 2;
}

// a host array into a device buffer through the context's pinned arena
template <class T>
static void stage_upload(Context &C, DBuf<T> &b, const T *src, size_t count) {
  b.alloc(count);
  C.h2d(b.p, src, count * sizeof(T));
}
template <class T>
static void stage_upload(Context &C, DBuf<T> &b, const std::vector<T> &v) {
  stage_upload(C, b, v.data(), v.size());
}

static double device_max(Context &C, const double *d, long long n, DBuf<double> &tmp) {
  if (n <= 0) return -INFINITY;
  int nb = (int)std::min<long long>(1024, (n + 255) / 256);
  tmp.alloc(1024);
  hipLaunchKernelGGL(k_blockmax, dim3(nb), dim3(256), 0, C.stream, d, n, tmp.p);
  std::vector<double> h(nb);
  C.read(h.data(), tmp.p, sizeof(double) * nb);
  double m = -INFINITY;
  for (double v : h) m = v > m ? v : m;
  return m;
}

struct EriPhaseTimer {
  Context &C;
  std::string name;
  double ms0;
  long long q0, f0;
  EriPhaseTimer(Context &c, const char *n) : C(c), name(n), ms0(c.eri_ms_acc), q0(c.eri_quartets_acc), f0(c.eri_flops_acc) {}
  ~EriPhaseTimer() {
    C.chol.eri_kernel_ms[name] += C.eri_ms_acc - ms0;
    C.chol.eri_quartets[name] += C.eri_quartets_acc - q0;
  }
};

// factor and invert the diagonal block [k0, k0 + m) of V (n x n); X (zero on
// entry) receives L^-1 in the same block; W: scratch of (m/2 + POTF_NB)^2
static void chol_inv_rec(Context &C, double *V, double *X, long long n, long long k0, long long m, double *W,
                         int *dinfo) {
  if (m <= POTF_NB) {
    hipLaunchKernelGGL(k_chol_inv_leaf, dim3(1), dim3(256), 0, C.stream, V, X, n, k0, (int)m, dinfo);
    return;
  }
  const long long m1 = (m / 2 + POTF_NB - 1) / POTF_NB * POTF_NB, m2 = m - m1;
  chol_inv_rec(C, V, X, n, k0, m1, W, dinfo);
  double *A21 = V + (k0 + m1) + n * k0, *A22 = V + (k0 + m1) + n * (k0 + m1);
  double *X11 = X + k0 + n * k0, *X21 = X + (k0 + m1) + n * k0, *X22 = X + (k0 + m1) + n * (k0 + m1);
  const double one = 1.0, mone = -1.0, zero = 0.0;
  // L21 = A21 X11^T, stored over A21
  ROCBLAS_CHECK(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_transpose, m2, m1, m1, &one, A21, n,
                              X11, n, &zero, W, m2));
  GI_HIP(hipMemcpy2DAsync(A21, sizeof(double) * n, W, sizeof(double) * m2, sizeof(double) * m2, m1,
                          hipMemcpyDeviceToDevice, C.stream));
  // A22 -= L21 L21^T (the whole square; only its lower triangle is read later)
  ROCBLAS_CHECK(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_transpose, m2, m2, m1, &mone, A21, n,
                              A21, n, &one, A22, n));
  chol_inv_rec(C, V, X, n, k0 + m1, m2, W, dinfo);
  // X21 = -X22 (L21 X11)
  ROCBLAS_CHECK(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_none, m2, m1, m1, &one, A21, n, X11,
                              n, &zero, W, m2));
  ROCBLAS_CHECK(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_none, m2, m1, m2, &mone, X22, n, W,
                              m2, &zero, X21, n));
}

// Cholesky factor of V (n x n, lower triangle read, L written over it) and its
// inverse into X (n x n, zeroed here); W: (n/2 + POTF_NB)^2 doubles, info: one
// device int.  Returns false when a pivot is not positive (V not numerically
// positive definite).  For callers outside the decomposition (the THC metric).
bool chol_factor_inverse(Context &C, double *V, double *X, long long n, double *W, int *info) {
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, C.stream, info, 0);
  GI_HIP(hipMemsetAsync(X, 0, sizeof(double) * (size_t)n * n, C.stream));
  chol_inv_rec(C, V, X, n, 0, n, W, info);
  int h = 0;
  C.read(&h, info, sizeof(h));
  return h == 0;
}
long long chol_factor_inverse_scratch(long long n) { return (n / 2 + POTF_NB) * (n / 2 + POTF_NB); }

// V (n x n, lower triangle significant) -> L^-1; throws the reference's error
// when V is not positive definite
static void factor_and_invert(Context &C, DBuf<double> &V, long long n) {
  CholWork &cw = chol_work(C);
  DBuf<int> &dinfo = cw.info;
  dinfo.alloc(1);
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, C.stream, dinfo.p, 0);
  DBuf<double> &X = cw.X, &W = cw.W;
  X.alloc((size_t)n * n);
  GI_HIP(hipMemsetAsync(X.p, 0, sizeof(double) * (size_t)n * n, C.stream));
  const long long h = n / 2 + POTF_NB;
  W.alloc((size_t)h * h);
  chol_inv_rec(C, V.p, X.p, n, 0, n, W.p, dinfo.p);
  int info = 0;
  C.read(&info, dinfo.p, sizeof(info));
  if (info != 0) throw Error{40, "Cholesky factorization failed with exit code info=" + std::to_string(info)};
  std::swap(V.p, X.p);
  std::swap(V.n, X.n);
  std::swap(V.st, X.st);
}

// r(pq) = Vdiag(full(pq)) - sum_k R(k, col)^2 for every Rkpq column; one
// wavefront per column (256 lanes = 4 columns per workgroup), padding columns
// (full index -1) give 0
__global__ void __launch_bounds__(256) k_residual_diag(const double *__restrict__ R, long long nvec,
                                                       const long long *__restrict__ colfull, long long ncol,
                                                       const double *__restrict__ vdiag, double *__restrict__ out) {
  const long long col = blockIdx.x * 4LL + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (col >= ncol) return;
  double s = 0.0;
  for (long long k = lane; k < nvec; k += 64) {
    const double v = R[k + nvec * col];
    s += v * v;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) out[col] = colfull[col] >= 0 ? vdiag[colfull[col]] - s : 0.0;
}

void chol_residual(Context &C, double *rmax, double *rmin) {
  const CholState &S = C.chol;
  if (!S.valid || !S.rkpq.p) throw Error{50, "Cholesky vectors are not available (run gi_chol_build first)"};
  const long long msd = S.max_subset_dim, ncol = msd * S.nsubsets;
  // Vdiag is stored shell pair by shell pair in chol2_Vdiag order (ShB <= ShA,
  // ShA fastest), each pair's orbital pairs in the order of its Rkpq columns
  const int nsh = C.nshells;
  std::vector<long long> vdloc((size_t)nsh * nsh, 0);
  for (long long sb = 0, loc = 0; sb < nsh; ++sb)
    for (long long sa = sb; sa < nsh; ++sa) {
      vdloc[sa * nsh + sb] = loc;
      loc += pair_dim_of(C, (int)sa, (int)sb);
    }
  std::vector<long long> colfull(ncol, -1);   // column -> Vdiag index
  for (long long s = 0; s < S.nshell_pairs; ++s) {
    const long long A = S.shell_pairs[2 * s] - 1, B = S.shell_pairs[2 * s + 1] - 1;
    const long long pos = S.shell_pair_loc[3 * s + 1] - 1, X = S.shell_pair_loc[3 * s + 2];
    for (long long q = 0; q < S.shell_pair_dim[s]; ++q) colfull[(X - 1) * msd + pos + q] = vdloc[A * nsh + B] + q;
  }
  DBuf<long long> dcol;
  DBuf<double> dv, dout;
  dcol.upload(colfull, C.stream);
  dv.upload(S.vdiag, C.stream);
  dout.alloc((size_t)ncol);
  hipLaunchKernelGGL(k_residual_diag, dim3((unsigned)((ncol + 3) / 4)), dim3(256), 0, C.stream, S.rkpq.p, S.nvecs,
                     dcol.p, ncol, dv.p, dout.p);
  GI_HIP(hipGetLastError());
  std::vector<double> r(ncol);
  GI_HIP(hipMemcpyAsync(r.data(), dout.p, sizeof(double) * ncol, hipMemcpyDeviceToHost, C.stream));
  C.wait();
  double mx = -INFINITY, mn = INFINITY;
  for (long long c = 0; c < ncol; ++c)
    if (colfull[c] >= 0) {
      mx = std::max(mx, r[c]);
      mn = std::min(mn, r[c]);
    }
  *rmax = mx;
  *rmin = mn;
}

// ------------------------------------------------------------------ driver
void chol_build(Context &C, double tau, double kappa, long long max_block_dim, long long block_dim,
                long long flags) {
  if (!C.basis_set) throw Error{4, "basis not set"};
  if (max_block_dim <= 0 || block_dim <= 0) throw Error{5, "MaxBlockDim and CholVecsBlock must be positive"};
  CholState &S = C.chol;
  S.valid = false;
  // Inv_L and Rkpq keep their device arrays across builds (grow-only); a
  // build without vectors must not leave an earlier build's Rkpq readable
  if (flags & 2) S.rkpq.release();
  S.mo_adj_off.release();
  S.mo_adj_col.release();
  S.mo_adj_p.release();
  S.timings.clear();
  S.eri_kernel_ms.clear();
  S.eri_quartets.clear();
  S.tau = tau;
  S.kappa = kappa;
  S.nmacro = 0;
  C.time_eri = (flags & 4) != 0;
  C.eri_ms_acc = 0;
  C.eri_quartets_acc = 0;
  C.eri_flops_acc = 0;
  C.eri_launches_acc = 0;
  const int nsh = C.nshells;
  const long long nao = C.nao;
  const long long max_sp = (long long)nsh * (nsh + 1) / 2;
  const long long max_op = nao * (nao + 1) / 2;
  const int maxnang = C.spher ? 2 * C.lmax + 1 : ncart(C.lmax);
  GI_HIP(hipSetDevice(C.device));
  const double t_begin = now();

  // ---------------- chol2_Vdiag
  std::vector<long long> vd_loc(max_sp + 1), vd_dim(max_sp + 1);
  std::vector<int> vd_a(max_sp + 1), vd_b(max_sp + 1);
  {
    long long loc = 1, ab = 1;
    for (int sb = 0; sb < nsh; ++sb)
      for (int sa = sb; sa < nsh; ++sa) {
        vd_loc[ab] = loc;
        vd_dim[ab] = pair_dim_of(C, sa, sb);
        vd_a[ab] = sa;
        vd_b[ab] = sb;
        loc += vd_dim[ab];
        ++ab;
      }
  }
  DBuf<double> d_vdiag;
  d_vdiag.alloc(max_op);
  {
    double t0 = now();
    EriPhaseTimer pt(C, "vdiag");
    std::vector<PairDesc> rd(max_sp), cd(max_sp);
    std::vector<int> rows(max_sp);
    for (long long k = 1; k <= max_sp; ++k) {
      int sa = vd_a[k], sb = vd_b[k];
      int fl = (sa == sb) ? PAIR_SAME : 0;
      rd[k - 1] = PairDesc{sa, sb, fl, 0, 0, 0, vd_loc[k] - 1};
      cd[k - 1] = PairDesc{sa, sb, fl, 0, 0, 0, 0};
      rows[k - 1] = (int)(k - 1);
    }
    EriRequest R;
    R.rdesc = &rd; R.cdesc = &cd;
    R.rows = &rows; R.cols = &rows; R.mode = LAUNCH_DIAG;
    R.out = d_vdiag.p; R.ldo = 0; R.kappa = kappa;
    eri_run(C, R);
    S.vdiag.resize(max_op);
    C.read(S.vdiag.data(), d_vdiag.p, sizeof(double) * max_op);
    S.timings["vdiag"] = now() - t0;
  }
  const std::vector<double> &vdiag = S.vdiag;

  // ---------------- chol2_DefineBase
  std::vector<double> vd_max(max_sp + 1);
  std::vector<long long> idxmap(max_sp + 1);
  for (long long k = 1; k <= max_sp; ++k) {
    double mx = -INFINITY;
    for (long long i = 0; i < vd_dim[k]; ++i) mx = std::max(mx, vdiag[vd_loc[k] - 1 + i]);
    vd_max[k] = mx;
    idxmap[k] = k;
  }
  double maxdrs = -INFINITY;
  for (long long k = 1; k <= max_sp; ++k) maxdrs = std::max(maxdrs, vd_max[k]);
  if (!(maxdrs > 0.0)) throw Error{10, "Maximum diagonal Coulomb integral is nonpositive"};
  double pthresh = std::min(tau, 1.0e-8);
  pthresh = pthresh * pthresh / maxdrs;
  const bool raw_ties = (flags & 8) != 0;
  int maxdrs_exp = 0;
  std::frexp(maxdrs, &maxdrs_exp);
  int tau_exp = 0;
  std::frexp(tau, &tau_exp);
  const double inv_delta = std::ldexp(1.0, std::max(KEY_ABS_BITS + 1 - maxdrs_exp, KEY_TAU_BITS + 1 - tau_exp));
  if (!raw_ties)
    for (long long k = 1; k <= max_sp; ++k) vd_max[k] = qrel(vd_max[k]);
  dsort0(vd_max.data() + 1, idxmap.data() + 1, max_sp, 2);
  long long k0 = 1;
  for (long long k = 1; k <= max_sp; ++k) {
    long long ab = idxmap[k];
    double dpq = -INFINITY;
    for (long long i = 0; i < vd_dim[ab]; ++i) dpq = std::max(dpq, vdiag[vd_loc[ab] - 1 + i]);
    if (dpq <= pthresh) k0 = k;   // the last discarded pair stays in BASE
    else break;
  }
  const long long nsp = max_sp - k0 + 1;
  long long nop = 0;
  for (long long k = k0; k <= max_sp; ++k) nop += vd_dim[idxmap[k]];
  // 1-based BASE tables
  std::vector<int> spA(nsp + 1), spB(nsp + 1);
  std::vector<long long> spdim(nsp + 1), loc_full(nsp + 1), loc_sub(nsp + 1, 0), idx_sub(nsp + 1, 0),
      loc_comp(nsp + 1);
  std::vector<double> Dh(nop);
  {
    long long s = 1, p = 1;
    for (long long k = k0; k <= max_sp; ++k) {
      long long l = idxmap[k];
      spA[s] = vd_a[l];
      spB[s] = vd_b[l];
      spdim[s] = vd_dim[l];
      loc_full[s] = p;
      for (long long i = 0; i < vd_dim[l]; ++i) Dh[p - 1 + i] = vdiag[vd_loc[l] - 1 + i];
      ++s;
      p += vd_dim[l];
    }
  }
  S.nshell_pairs = nsp;
  S.norb_pairs = nop;

  // ---------------- chol2_DefineSubsets / DivideOrbPairs (one image)
  const long long nsub = (long long)std::ceil((double)nop / (double)max_block_dim);
  std::vector<long long> sdim(nsub + 1, 0), sb0(nsub + 1, 0), sb1(nsub + 1, 0);
  {
    long long target = (long long)std::ceil((double)nop / (double)nsub);
    long long berr[2] = {0, 0}, bloc[2] = {0, 0};
    for (long long k = 1; k <= nsub; ++k) {
      if (k == 1) sb0[k] = 1;
      else sb0[k] = (sb1[k - 1] < nsp && sb1[k - 1] > 0) ? sb1[k - 1] + 1 : 0;
      sdim[k] = 0;
      sb1[k] = sb0[k];
      if (sb0[k] > 0) {
        if (k < nsub) {
          berr[0] = nop;
          berr[1] = nop;
          for (long long s = sb0[k]; s <= nsp; ++s) {
            loc_sub[s] = sdim[k] + 1;
            idx_sub[s] = k;
            sdim[k] += spdim[s];
            if (sdim[k] >= target) { bloc[1] = s; berr[1] = sdim[k] - target; break; }
            bloc[0] = s;
            berr[0] = sdim[k] - target;
          }
          if (std::llabs(berr[0]) < std::llabs(berr[1])) { sb1[k] = bloc[0]; sdim[k] = berr[0] + target; }
          else { sb1[k] = bloc[1]; sdim[k] = berr[1] + target; }
        } else {
          sb1[k] = nsp;
          for (long long s = sb0[k]; s <= nsp; ++s) {
            loc_sub[s] = sdim[k] + 1;
            idx_sub[s] = k;
            sdim[k] += spdim[s];
          }
        }
      }
    }
  }
  long long msd = 0;
  for (long long k = 1; k <= nsub; ++k) msd = std::max(msd, sdim[k]);
  S.nsubsets = nsub;
  S.max_subset_dim = msd;
  const long long max_nchol = (long long)std::min(std::ceil((double)nao * 15.0), (double)nop);

  // ---------------- chol2_Pivots
  std::vector<long long> pivots(nop + 1, 0), pivot_order;
  long long nchol = 0;
  bool converged = false;
  double dmax = 0.0;
  double t_eri1 = 0, t_rtr = 0, t_micro = 0, rtr_flops = 0, t_cand = 0, t_batch = 0;
  {
    double t0 = now();
    long long nsps_cand = nsp, nops_cand = nop;
    std::vector<long long> cand(nsp + 1);   // candidate shell pairs, compressed order
    for (long long k = 1; k <= nsp; ++k) { cand[k] = k; loc_comp[k] = loc_full[k]; }
    CholWork &cw = chol_work(C);
    DBuf<double> &dD = cw.D, &dDn = cw.Dn, &dL = cw.L, &dLn = cw.Ln, &dM = cw.M, &dLB = cw.LB, &dtmp = cw.tmp,
                 &dbm = cw.bm, &dMb = cw.Mb, &dLb = cw.Lb, &dDb = cw.Db, &dLP = cw.LP;
    DBuf<int> &dpiv = cw.piv;
    DBuf<MicroState> &dstate = cw.state;
    stage_upload(C, dD, Dh.data(), (size_t)nop);
    long long ncand = nop;       // rows of L / length of D
    long long capL = 0;          // columns allocated in L
    dbm.alloc(2);
    dmax = -INFINITY;
    for (double v : Dh) dmax = std::max(dmax, v);
    bool do_macro = max_nchol > 0;
    DBuf<long long> &dsrc = cw.src, &dloc = cw.loc, &dbd = cw.bd;
    DBuf<unsigned char> &dexcl = cw.excl;
    DBuf<int> &ddim = cw.dim;
    std::vector<double> mv;
    std::vector<long long> im;
    std::vector<double> spmax(nsp + 1, 0.0);   // raw residual maximum per BASE shell pair
    while (do_macro) {
      S.nmacro++;
      double tc0 = now();
      // chol2_DefineCandidates: max D per candidate pair, descending sort
      long long n = nsps_cand;
      mv.assign(n + 1, 0.0);
      im.assign(n + 1, 0);
      if (n > 0) {
        std::vector<long long> hl(n);
        std::vector<int> hd(n);
        for (long long k = 1; k <= n; ++k) { hl[k - 1] = loc_comp[cand[k]] - 1; hd[k - 1] = (int)spdim[cand[k]]; }
        stage_upload(C, dloc, hl);
        stage_upload(C, ddim, hd);
        dtmp.alloc(n);
        hipLaunchKernelGGL(k_segmax, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, C.stream, dD.p, dloc.p,
                           ddim.p, n, dtmp.p);
        C.read(mv.data() + 1, dtmp.p, sizeof(double) * n);
        for (long long k = 1; k <= n; ++k) {
          im[k] = cand[k];
          spmax[cand[k]] = mv[k];
          if (!raw_ties) mv[k] = qabs(mv[k], inv_delta) / inv_delta;
        }
      }
      // sorted by the (quantized) key; the threshold compares the raw maxima,
      // which within one quantum need not be monotone, so every pair is tested
      // (with raw keys this is the reference's loop up to its first miss)
      dsort0(mv.data() + 1, im.data() + 1, n, -2);
      nsps_cand = 0;
      nops_cand = 0;
      for (long long k = 1; k <= n; ++k) {
        if (spmax[im[k]] > tau) {
          ++nsps_cand;
          cand[nsps_cand] = im[k];
          nops_cand += spdim[im[k]];
        } else if (raw_ties) break;
      }
      // chol2_Compress: move D and the stored vectors to the new candidate order
      long long max_nbatch_cap = std::max<long long>((long long)maxnang * maxnang, 1000);
      if (nsps_cand > 0) {
        std::vector<long long> src;
        src.reserve(nops_cand);
        std::vector<long long> newloc(nsps_cand + 1);
        long long ncomp = 0;
        for (long long g = 1; g <= nsps_cand; ++g) {
          long long ab = cand[g];
          newloc[g] = ncomp + 1;
          for (long long i = 0; i < spdim[ab]; ++i) src.push_back(loc_comp[ab] - 1 + i);
          ncomp += spdim[ab];
        }
        stage_upload(C, dsrc, src);
        dDn.alloc(ncomp);
        hipLaunchKernelGGL(k_gather_rows_of_cols, dim3((unsigned)((ncomp + 255) / 256), 1), dim3(256), 0, C.stream,
                           dDn.p, ncomp, dD.p, ncand, dsrc.p, ncomp, 1LL);
        std::swap(dD.p, dDn.p);
        std::swap(dD.n, dDn.n);
        std::swap(dD.st, dDn.st);
        // new L with room for this macro iteration's vectors
        long long mnb = std::min(ncomp, max_nbatch_cap);
        long long newcap = std::min(max_nchol, nchol + mnb);
        newcap = std::max(newcap, (long long)1);
        dLn.alloc((size_t)ncomp * newcap);
        if (nchol > 0) {
          unsigned gy = (unsigned)std::min<long long>(nchol, 65535);
          hipLaunchKernelGGL(k_gather_rows_of_cols, dim3((unsigned)((ncomp + 255) / 256), gy), dim3(256), 0,
                             C.stream, dLn.p, ncomp, dL.p, ncand, dsrc.p, ncomp, nchol);
        }
        std::swap(dL.p, dLn.p);
        std::swap(dL.n, dLn.n);
        std::swap(dL.st, dLn.st);
        capL = newcap;
        for (long long g = 1; g <= nsps_cand; ++g) loc_comp[cand[g]] = newloc[g];
        ncand = ncomp;
      }
      t_cand += now() - tc0;
      double tb0 = now();
      const long long max_nbatch = std::min(nops_cand, max_nbatch_cap);
      double dmin = std::max(0.01 * dmax, tau);
      // chol2_DefineBatch: candidates are already in descending order of their
      // maxima; the reference sorts them again (ties may be permuted)
      std::vector<double> ds(nsps_cand + 1);
      std::vector<long long> di(nsps_cand + 1);
      for (long long k = 1; k <= nsps_cand; ++k) {
        di[k] = cand[k];
        ds[k] = raw_ties ? spmax[cand[k]] : qabs(spmax[cand[k]], inv_delta) / inv_delta;
      }
      dsort0(ds.data() + 1, di.data() + 1, nsps_cand, -2);
      long long nsps_batch = 0, nops_batch = 0;
      std::vector<long long> bsp(1), colloc(1);
      for (long long k = 1; k <= nsps_cand; ++k) {
        long long ab = di[k];
        if (spmax[ab] > dmin && nops_batch + spdim[ab] <= max_nbatch) {
          bsp.push_back(ab);
          colloc.push_back(nops_batch + 1);
          nsps_batch = k;
          nops_batch += spdim[ab];
        } else break;
      }
      // batch element tables: compressed row, FULL location
      std::vector<long long> bd(nops_batch), bfull(nops_batch);
      for (long long k = 1; k <= nsps_batch; ++k)
        for (long long p = 0; p < spdim[bsp[k]]; ++p) {
          bd[colloc[k] - 1 + p] = loc_comp[bsp[k]] - 1 + p;
          bfull[colloc[k] - 1 + p] = loc_full[bsp[k]] + p;
        }
      stage_upload(C, dbd, bd);
      t_batch += now() - tb0;
      // chol2_M: ERIs (candidates | batch)
      dM.alloc((size_t)std::max<long long>(ncand, 1) * std::max<long long>(max_nbatch, 1));
      {
        double te = now();
        EriPhaseTimer pt(C, "step1");
        std::vector<PairDesc> rd(nsps_cand), cd(nsps_batch);
        std::vector<int> rows(nsps_cand), cols(nsps_batch);
        for (long long k = 1; k <= nsps_cand; ++k) {
          long long ab = cand[k];
          rd[k - 1] = PairDesc{spA[ab], spB[ab], spA[ab] == spB[ab] ? PAIR_SAME : 0, 0, 0, 0, loc_comp[ab] - 1};
          rows[k - 1] = (int)(k - 1);
        }
        for (long long k = 1; k <= nsps_batch; ++k) {
          long long ab = bsp[k];
          cd[k - 1] = PairDesc{spA[ab], spB[ab], spA[ab] == spB[ab] ? PAIR_SAME : 0, 0, 0, 0, ncand * (colloc[k] - 1)};
          cols[k - 1] = (int)(k - 1);
        }
        if (nsps_batch > 0 && nsps_cand > 0) {
          EriRequest R;
          R.rdesc = &rd; R.cdesc = &cd;
          R.rows = &rows; R.cols = &cols; R.mode = LAUNCH_PRODUCT;
          R.out = dM.p; R.ldo = ncand; R.kappa = kappa;
          eri_run(C, R);
        }
        t_eri1 += now() - te;
      }
      // chol2_SubtractR: M(:, batch) -= L^T L(batch)  (vectors are columns of L)
      if (nchol > 0 && nops_batch > 0) {
        double tr = now();
        dLB.alloc((size_t)nops_batch * nchol);
        unsigned gy = (unsigned)std::min<long long>(nchol, 65535);
        hipLaunchKernelGGL(k_gather_rows_of_cols, dim3((unsigned)((nops_batch + 255) / 256), gy), dim3(256), 0,
                           C.stream, dLB.p, nops_batch, dL.p, ncand, dbd.p, nops_batch, nchol);
        const double alpha = -1.0, beta = 1.0;
        if (flags & 16)
          split_bf16_gemm(C, rocblas_operation_none, rocblas_operation_transpose, ncand, nops_batch, nchol, alpha,
                          dL.p, ncand, dLB.p, nops_batch, beta, dM.p, ncand);
        else
          ROCBLAS_CHECK(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_transpose, ncand,
                                      nops_batch, nchol, &alpha, dL.p, ncand, dLB.p, nops_batch, &beta, dM.p, ncand));
        C.wait();
        t_rtr += now() - tr;
        rtr_flops += 2.0 * (double)ncand * (double)nops_batch * (double)nchol;
      }
      // micro iterations (chol2_Lj + D update + BatchMax, TwoStepCholesky_Step1.f90:1027-1064):
      // the sequential recurrence runs on the batch rows on the device; the
      // other candidate rows follow from one triangular solve
      //   L(:, j0:j0+J) = M(:, P) L_P^-T,  L_P(q, r) = L(p_q, r),
      // which is the same forward substitution row by row.
      double tm = now();
      const long long j0 = nchol;
      const long long nb = nops_batch;
      const int nsteps = (int)std::min<long long>(nb, max_nchol - nchol);
      long long J = 0;
      if (nchol < max_nchol && nb > 0) {
        dMb.alloc((size_t)nb * nb);
        dLb.alloc((size_t)nb * std::max(nsteps, 1));
        dDb.alloc(nb);
        dpiv.alloc(nb);
        dstate.alloc(1);
        if (nb > MICRO_ROWS) throw Error{21, "batch larger than the micro-iteration workgroup"};
        hipLaunchKernelGGL(k_gather_block, dim3((unsigned)((nb + 255) / 256), (unsigned)nb), dim3(256), 0, C.stream,
                           dMb.p, dM.p, ncand, dbd.p, nb);
        const double qd = raw_ties ? 0.0 : inv_delta;
        const bool split = (flags & 16) != 0;
        if (split) {
          std::vector<unsigned char> ex(nb);
          for (long long e = 0; e < nb; ++e) ex[e] = pivots[bfull[e]] ? 1 : 0;
          stage_upload(C, dexcl, ex);
        }
        hipLaunchKernelGGL(k_micro_init, dim3(1), dim3(1024), 0, C.stream, dD.p, dbd.p, nb, qd, dDb.p, dstate.p,
                           nsteps, split ? dMb.p : nullptr, split ? dexcl.p : nullptr);
        // panels until the loop stops: in one launch for batch blocks up to
        // micro_persist_max rows, else a panel and an update launch per
        // MICRO_PANEL steps with the state read back every 4 panels
        const int npanel_max = (nsteps + MICRO_PANEL - 1) / MICRO_PANEL;
        if (nb <= C.micro_persist_max) {
          hipLaunchKernelGGL(k_micro_all, dim3(1), dim3(MICRO_ROWS), 0, C.stream, dstate.p, dMb.p, dLb.p, dDb.p, nb,
                             nsteps, dmin, qd, dpiv.p);
          GI_HIP(hipGetLastError());
        }
        for (int pnl = nb <= C.micro_persist_max ? npanel_max : 0; pnl < npanel_max;) {
          for (int k = 0; k < 4 && pnl < npanel_max; ++k, ++pnl) {
            hipLaunchKernelGGL(k_micro_panel, dim3(1), dim3(MICRO_ROWS), 0, C.stream, dstate.p, dMb.p, dLb.p,
                               dDb.p, nb, nsteps, dmin, qd, dpiv.p);
            hipLaunchKernelGGL(k_micro_update, dim3((unsigned)((nb + 255) / 256), (unsigned)nb), dim3(256), 0,
                               C.stream, dstate.p, dMb.p, dLb.p, nb);
          }
          MicroState hs{};
          C.read(&hs, dstate.p, sizeof(hs));
          if (hs.done) break;
        }
        MicroState hs{};
        C.read(&hs, dstate.p, sizeof(hs));
        J = hs.j;
        std::vector<int> hpiv(std::max<long long>(J, 1));
        if (J > 0) {
          C.read(hpiv.data(), dpiv.p, sizeof(int) * J);
          for (long long q = 0; q < J; ++q) {
            pivots[bfull[hpiv[q]]] = 1;
            pivot_order.push_back(bfull[hpiv[q]]);
          }
          // all candidate rows: right-hand sides, L_P, triangular solve
          hipLaunchKernelGGL(k_gather_cols, dim3((unsigned)((ncand + 255) / 256), (unsigned)J), dim3(256), 0,
                             C.stream, dL.p + (size_t)ncand * j0, ncand, dM.p, ncand, dpiv.p, ncand);
          dLP.alloc((size_t)J * J);
          hipLaunchKernelGGL(k_build_lp, dim3((unsigned)((J + 255) / 256), (unsigned)J), dim3(256), 0, C.stream,
                             dLP.p, J, dLb.p, nb, dpiv.p);
          const double one = 1.0;
          ROCBLAS_CHECK(rocblas_dtrsm(C.blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                                      rocblas_diagonal_non_unit, ncand, J, &one, dLP.p, J,
                                      dL.p + (size_t)ncand * j0, ncand));
          hipLaunchKernelGGL(k_finish_rows, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, C.stream, dL.p, ncand,
                             j0, J, dLb.p, nb, dbd.p, dD.p, ncand);
          hipLaunchKernelGGL(k_update_d, dim3((unsigned)((ncand + 255) / 256)), dim3(256), 0, C.stream, dL.p,
                             ncand, j0, J, dD.p, ncand);
          if (split)
            hipLaunchKernelGGL(k_zero_pivot_d, dim3((unsigned)((J + 255) / 256)), dim3(256), 0, C.stream, dD.p,
                               dbd.p, dpiv.p, J);
        }
        nchol += J;
      }
      t_micro += now() - tm;
      // maximum of the raw residual diagonal over the rows still held (with no
      // candidate left these are the previous candidates, all <= tau)
      dmax = (ncand > 0) ? device_max(C, dD.p, ncand, dtmp) : 0.0;
      if (dmax <= tau) { converged = true; do_macro = false; }
      if (nchol == max_nchol || nops_batch == 0) do_macro = false;
    }
    S.timings["step1"] = now() - t0;
  }
  S.timings["eri_step1"] = t_eri1;
  S.timings["rtr"] = t_rtr;
  S.timings["rtr_flops"] = rtr_flops;
  S.timings["micro"] = t_micro;
  S.timings["candidates"] = t_cand;
  S.timings["batch"] = t_batch;
  S.max_residual = dmax;
  S.nvecs = nchol;
  // step 1 outputs
  S.shell_pairs.assign(2 * nsp, 0);
  S.shell_pair_loc.assign(3 * nsp, 0);
  S.shell_pair_dim.assign(nsp, 0);
  for (long long s = 1; s <= nsp; ++s) {
    S.shell_pairs[2 * (s - 1)] = spA[s] + 1;
    S.shell_pairs[2 * (s - 1) + 1] = spB[s] + 1;
    S.shell_pair_loc[3 * (s - 1)] = loc_full[s];
    S.shell_pair_loc[3 * (s - 1) + 1] = loc_sub[s];
    S.shell_pair_loc[3 * (s - 1) + 2] = idx_sub[s];
    S.shell_pair_dim[s - 1] = spdim[s];
  }
  S.subset_dim.assign(nsub, 0);
  S.subset_bounds.assign(2 * nsub, 0);
  for (long long k = 1; k <= nsub; ++k) {
    S.subset_dim[k - 1] = sdim[k];
    S.subset_bounds[2 * (k - 1)] = sb0[k];
    S.subset_bounds[2 * (k - 1) + 1] = sb1[k];
  }
  S.pivots.assign(pivots.begin() + 1, pivots.end());
  S.pivot_order = pivot_order;
  if (nchol == max_nchol && !converged)
    throw Error{20, "Max number of Cholesky vectors reached without convergence"};

  // ---------------- chol2_Step2
  const long long nvec = nchol;
  double t2 = now();
  std::vector<long long> psp, psploc, pspdim, porb;
  {
    long long nsum = 0;
    for (long long i = 1; i <= nop; ++i) nsum += pivots[i];
    if (nsum != nvec) throw Error{30, "Invalid Pivots array"};
    long long q = 0;
    for (long long ab = 1; ab <= nsp; ++ab) {
      long long p0 = loc_full[ab], p1 = p0 + spdim[ab] - 1, s = 0;
      for (long long p = p0; p <= p1; ++p) s += pivots[p];
      if (s == 0) continue;
      psp.push_back(ab);
      psploc.push_back(q + 1);
      pspdim.push_back(s);
      if (spA[ab] != spB[ab]) {
        for (long long p = p0; p <= p1; ++p)
          if (pivots[p] == 1) { porb.push_back(p - p0 + 1); ++q; }
      } else {
        long long na = C.shell_nang[spA[ab]];
        for (long long p = p0; p <= p1; ++p)
          if (pivots[p] == 1) {
            long long pos = p - p0 + 1, a = 0, b = 0;
            if (flags & 1) {
              long long cnt = 0;
              for (long long y = 1; y <= na && !a; ++y)
                for (long long x = y; x <= na; ++x)
                  if (++cnt == pos) { a = x; b = y; break; }
            } else {
              // chol2_pq2p_ge_q, TwoStepCholesky_Step2.f90:422-466
              long long qb = (pos - 1) / (na + 1) + 1;
              long long v = pos - (na + 1) * (qb - 1);
              long long i1 = na - qb + 1;
              long long in2 = v / (i1 + 1), in1 = in2 ^ 1;
              a = in1 * (qb + v - 1) + in2 * (v - i1 + na - qb);
              b = in1 * qb + in2 * i1;
            }
            porb.push_back(a + (b - 1) * na);
            ++q;
          }
      }
    }
  }
  const long long npsp = (long long)psp.size();
  S.npivot_shell_pairs = npsp;
  S.pivot_shell_pairs = psp;
  S.pivot_shell_pair_loc = psploc;
  S.pivot_shell_pair_dim = pspdim;
  S.pivot_orb_pairs = porb;
  // selection list: 0-based full-block indices of the pivots
  std::vector<int> sel(nvec);
  for (long long i = 0; i < nvec; ++i) sel[i] = (int)(porb[i] - 1);
  DBuf<int> dsel;
  stage_upload(C, dsel, sel);
  // chol2_Vpqrs
  S.inv_l.alloc((size_t)nvec * nvec);
  {
    double tv = now();
    EriPhaseTimer pt(C, "vpqrs");
    std::vector<PairDesc> rd(npsp), cd(npsp);
    std::vector<int> rows(npsp);
    for (long long k = 0; k < npsp; ++k) {
      long long ab = psp[k];
      rd[k] = PairDesc{spA[ab], spB[ab], PAIR_SEL, (int)(psploc[k] - 1), (int)pspdim[k], 0, psploc[k] - 1};
      cd[k] = PairDesc{spA[ab], spB[ab], PAIR_SEL, (int)(psploc[k] - 1), (int)pspdim[k], 0, nvec * (psploc[k] - 1)};
      rows[k] = (int)k;
    }
    EriRequest R;
    R.rdesc = &rd; R.cdesc = &cd;
    R.rows = &rows; R.cols = &rows; R.mode = LAUNCH_PRODUCT;
    R.out = S.inv_l.p; R.ldo = nvec; R.d_sel = dsel.p; R.kappa = kappa;
    eri_run(C, R);
    S.timings["vpqrs"] = now() - tv;
  }
  // V = L L^T, L^-1 (real_Cholesky + real_LowerTriangularInverse)
  {
    double tc = now();
    factor_and_invert(C, S.inv_l, nvec);
    S.timings["potrf_trtri"] = now() - tc;
  }
  S.timings["step2"] = now() - t2;

  // ---------------- full-dimension vectors, Rkpq(:,:,X) = Inv_L Vabrs^T
  if (!(flags & 2)) {
    double tf = now();
    const size_t slab = (size_t)msd * nvec;
    S.rkpq.alloc((size_t)nvec * msd * nsub);
    // subsets per group: keep the Vabrs staging under ~16 GiB
    long long per_group = std::max<long long>(1, (long long)((16.0 * (1ull << 30)) / (8.0 * (double)std::max<size_t>(slab, 1))));
    per_group = std::min(per_group, nsub);
    DBuf<double> &dW = chol_work(C).Wv;
    dW.alloc(slab * per_group);
    DBuf<long long> gpad;   // subset dimensions of the group
    double t_vab = 0, t_gemm = 0;
    for (long long X0 = 1; X0 <= nsub; X0 += per_group) {
      long long X1 = std::min(nsub, X0 + per_group - 1);
      long long ng = X1 - X0 + 1;
      // Vabrs of the group's subsets stacked into one (ng msd) x NVecs matrix:
      // row (X - X0) msd + subset position, column = pivot
      const long long ldw = msd * ng;
      // the ERI blocks below cover rows 0 .. sdim[X] - 1 of each subset and
      // every pivot column; only the padding rows up to msd need zeros
      {
        std::vector<long long> gdim(ng);
        for (long long X = X0; X <= X1; ++X) gdim[X - X0] = sdim[X];
        stage_upload(C, gpad, gdim);
        hipLaunchKernelGGL(k_zero_pad_rows, dim3((unsigned)std::min<long long>(nvec, 1024), (unsigned)ng), dim3(64), 0,
                           C.stream, dW.p, ldw, msd, gpad.p, nvec);
        GI_HIP(hipGetLastError());
      }
      std::vector<PairDesc> cd(npsp);
      std::vector<int> cols(npsp);
      for (long long k = 0; k < npsp; ++k) {
        long long ab = psp[k];
        cd[k] = PairDesc{spA[ab], spB[ab], PAIR_SEL, (int)(psploc[k] - 1), (int)pspdim[k], 0, ldw * (psploc[k] - 1)};
        cols[k] = (int)k;
      }
      std::vector<PairDesc> rd;
      std::vector<int> rows;
      for (long long X = X0; X <= X1; ++X) {
        if (sb0[X] <= 0) continue;
        for (long long s = sb0[X]; s <= sb1[X]; ++s) {
          rd.push_back(PairDesc{spA[s], spB[s], spA[s] == spB[s] ? PAIR_SAME : 0, 0, 0, 0,
                                (X - X0) * msd + loc_sub[s] - 1});
          rows.push_back((int)rd.size() - 1);
        }
      }
      double tv = now();
      if (!rd.empty() && npsp > 0) {
        EriPhaseTimer pt(C, "vabrs");
        EriRequest R;
        R.rdesc = &rd; R.cdesc = &cd;
        R.rows = &rows; R.cols = &cols; R.mode = LAUNCH_PRODUCT;
        R.out = dW.p; R.ldo = ldw; R.d_sel = dsel.p; R.kappa = kappa;
        eri_run(C, R);
      }
      t_vab += now() - tv;
      double tg = now();
      // Rkpq(:, :, X0..X1) is one NVecs x (ng msd) matrix (subsets are
      // consecutive in Rkpq), = Inv_L Vabrs^T.  Inv_L is lower triangular:
      // block row r only needs the first r1 columns of Inv_L, so the product
      // is one GEMM per block row ((nblk + 1) / (2 nblk) of the dense FLOPs)
      const double one = 1.0, zero = 0.0;
      // Block rows hold a multiple of 128 rows (the DGEMM macro tile), counted
      // from the last row: only the first block, whose product is the
      // shortest (k = r1), has a partly filled tile row.
      const long long nblk = std::max<long long>(1, std::min<long long>(C.vec_blocks, nvec / 256));
      const long long rows_blk = std::max<long long>(128, (nvec + 64 * nblk) / (128 * nblk) * 128);
      for (long long r1 = nvec; r1 > 0;) {
        const long long r0 = std::max<long long>(0, r1 - rows_blk);
        ROCBLAS_CHECK(rocblas_dgemm(C.blas, rocblas_operation_none, rocblas_operation_transpose, r1 - r0, ldw, r1,
                                    &one, S.inv_l.p + r0, nvec, dW.p, ldw, &zero,
                                    S.rkpq.p + (size_t)(X0 - 1) * nvec * msd + r0, nvec));
        r1 = r0;
      }
      C.wait();
      t_gemm += now() - tg;
    }
    S.timings["vabrs"] = t_vab;
    S.timings["vectors_gemm"] = t_gemm;
    S.timings["vectors"] = now() - tf;
  }
  C.wait();
  S.timings["total"] = now() - t_begin;
  S.nquartets = C.eri_quartets_acc;
  S.eri_flops = C.eri_flops_acc;
  S.timings["eri_launches"] = (double)C.eri_launches_acc;
  S.timings["eri_kernel"] = C.eri_ms_acc / 1000.0;
  S.valid = true;
}

}  // namespace gi
