// Split-BF16 products on MFMA: the precision form of the library's FP64
// GEMMs (the Cholesky trailing update and the AO->MO transform).
//
// Error-free splitting (Ozaki, Ogita, Oishi, Rump, Numer. Algorithms 59
// (2012) 95-118) onto BF16 MFMA with FP32 accumulation, k taken in chunks of
// 256:
//  * within a chunk, row r of op(A) is scaled by 2^-e and column c of op(B)
//    by 2^-f, the powers of two that bring the largest element of the row
//    (column) in that chunk into [1/2, 1);
//  * a scaled element x is cut into four slices of eight bits,
//      x = sum_t q_t 2^(-8t) + d,  |d| <= 2^-33,
//    q_1 in [-256, 256] and q_t in [-128, 128] integers, exact in BF16;
//  * A B is the sum over slice pairs (t, u) with t + u <= 5 (ten products,
//    the omitted pairs are below 2^-33) of 2^(-8(t+u)) q_t p_u; the pairs of
//    one level t + u share an FP32 accumulator;
//  * the leading levels carry at most 2^16 per k (q_1 p_1, and q_1 p_2 +
//    q_2 p_1), so their partial sums over the 256 k of a chunk are integers
//    of at most 2^24: exact in FP32 whatever the order of the MFMA's
//    additions.  Levels 4 and 5 are below 2^-17 of the product scale; their
//    FP32 rounding (2^-24 of that) stays under 2^-41.  At the end of a chunk
//    the level sums are scaled by 2^(e + f - 8 level) and added into an FP64
//    accumulator.
// The only error is the truncation of the slices: per chunk
//   |err(r, c)| <~ 2^-32 sum_k (max|A(r,chunk)| |B(k,c)| + |A(r,k)| max|B(chunk,c)|),
// with no growth from the accumulation over k (plain FP32 accumulation adds
// ~2^-24 sqrt(k) sum_k |A(r,k)||B(k,c)|).
#include "gi_context.h"

namespace gi {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int SPL_S = 4;      // slices per element
constexpr int SPL_KC = 256;   // k per FP32 chunk; k is padded to a multiple
constexpr int SPL_BM = 64;    // workgroup tile rows: 2 waves x 32
constexpr int SPL_BN = 128;   // workgroup tile columns: 2 waves x 64
constexpr int SPL_T = 32;     // slicing tile: 32 rows x one chunk of k

// Slices of the rows r < rows_pad, k < kpad (zero outside rows x kd; element
// (r, k) at X[r rs + k ks], one of rs, ks is 1) in the MFMA fragment order:
// unit = 8 BF16 (16 bytes), unit index
//   ((t nrb + rb) nks + ks) 64 + l,   rb = r / 16, ks = k / 32,
//   l = r % 16 + 16 ((k % 32) / 8),
// so one wave reads the 16 x 32 fragment (rb, ks) of slice t as one
// contiguous kilobyte.  e[ci rows_pad + r] = the scale exponent of row r in
// chunk ci.  One workgroup per 32 rows x 256 k tile, read along the
// contiguous index into LDS.
__global__ __launch_bounds__(256) void k_slice_rows(const double *__restrict__ X, long long rs, long long ks,
                                                    long long rows, long long kd, long long rows_pad, long long kpad,
                                                    bf16x8 *__restrict__ P, int *__restrict__ e) {
  __shared__ double t[SPL_T][SPL_KC + 1];
  __shared__ int se[SPL_T];
  const long long r0 = (long long)blockIdx.x * SPL_T, ci = blockIdx.y, k0 = ci * SPL_KC;
  const bool rfast = rs == 1;
  for (int i = threadIdx.x; i < SPL_T * SPL_KC; i += blockDim.x) {
    int ri, ki;
    if (rfast) {
      ri = i % SPL_T;
      ki = i / SPL_T;
    } else {
      ki = i % SPL_KC;
      ri = i / SPL_KC;
    }
    const long long r = r0 + ri, k = k0 + ki;
    t[ri][ki] = (r < rows && k < kd) ? X[r * rs + k * ks] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x < SPL_T) {
    double m = 0.0;
    for (int i = 0; i < SPL_KC; ++i) m = fmax(m, fabs(t[threadIdx.x][i]));
    int ex = 0;
    if (m > 0.0) (void)frexp(m, &ex);
    se[threadIdx.x] = ex;
    e[ci * rows_pad + r0 + threadIdx.x] = ex;
  }
  __syncthreads();
  const long long nrb = rows_pad / 16, nks = kpad / 32;
  const size_t plane = (size_t)nrb * (size_t)nks * 64;
  // 2 row blocks x 8 k steps x 64 lanes of 8 k each in the tile
  for (int i = threadIdx.x; i < (SPL_T / 16) * (SPL_KC / 32) * 64; i += blockDim.x) {
    const int l = i & 63, kb = (i >> 6) % (SPL_KC / 32), rbl = (i >> 6) / (SPL_KC / 32);
    const int ri = 16 * rbl + (l & 15), kq = 32 * kb + 8 * (l >> 4);
    const size_t u = ((size_t)(r0 / 16 + rbl) * (size_t)nks + (size_t)(k0 / 32 + kb)) * 64 + (size_t)l;
    double a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = ldexp(t[ri][kq + j], 8 - se[ri]);   // |a| < 256
#pragma unroll
    for (int s = 0; s < SPL_S; ++s) {
      bf16x8 v;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const double q = rint(a[j]);
        v[j] = (__bf16)(float)q;
        a[j] = (a[j] - q) * 256.0;                 // exact: |a - q| <= 1/2
      }
      P[s * plane + u] = v;
    }
  }
}

// C(r, c) = beta C + alpha sum_chunks 2^(ea + eb) sum_k x_r(k) y_c(k), x and y
// the slice sums of the rows of A (mpad rows) and B (npad rows), both in the
// fragment order of k_slice_rows, ea and eb their chunk exponents.  Workgroup: 4 waves, tile 64 x 128 (4 row
// and 8 column blocks of 16); wave: 32 x 64 as 2 x 4 MFMA 16x16x32 blocks.
// The 48 fragments of a 32-k step (4 + 8 blocks x 4 slices, 1 KB each) are
// copied global -> LDS by the four waves (global_load_lds, 16 bytes per
// lane, no registers), double-buffered: step s + 1 loads while step s
// multiplies.  Lane l feeds row (column) l % 16 of a block and the eight k
// at 8 (l / 16) of the step; it holds rows 4 (l / 16) + v, column l % 16 of
// the 16 x 16 result.  Every load is inside the padded planes; only the
// stores are bounded by m x n.
constexpr int SPL_NF = (SPL_BM / 16 + SPL_BN / 16) * SPL_S;   // fragments per k step

__device__ static inline void stage_step(bf16x8 (*st)[64], const bf16x8 *__restrict__ A,
                                         const bf16x8 *__restrict__ B, size_t pa, size_t pb, long long rb0,
                                         long long cb0, long long nks, long long kk, int w, int lane) {
#pragma unroll
  for (int q = 0; q < SPL_NF / 4; ++q) {
    const int f = w * (SPL_NF / 4) + q;
    const bf16x8 *src;
    if (f < (SPL_BM / 16) * SPL_S) {
      const int rb = f / SPL_S, t = f % SPL_S;
      src = A + t * pa + ((size_t)(rb0 + rb) * nks + kk) * 64 + lane;
    } else {
      const int g = f - (SPL_BM / 16) * SPL_S, cb = g / SPL_S, t = g % SPL_S;
      src = B + t * pb + ((size_t)(cb0 + cb) * nks + kk) * 64 + lane;
    }
    __builtin_amdgcn_global_load_lds((const void *)src, (void __attribute__((address_space(3))) *)(&st[f][0]), 16, 0,
                                     0);
  }
}

__global__ __launch_bounds__(256) void k_split_gemm(const bf16x8 *__restrict__ A, const bf16x8 *__restrict__ B,
                                                    const int *__restrict__ ea, const int *__restrict__ eb,
                                                    long long mpad, long long npad, long long kpad, long long m,
                                                    long long n, double alpha, double beta, double *__restrict__ C,
                                                    long long ldc) {
  __shared__ bf16x8 lds[2][SPL_NF][64];
  // XCD-aware tile order: workgroups go round-robin to the 8 XCDs (each with
  // its own L2), so workgroup b runs on XCD b % 8.  The column tiles of one
  // row tile are given consecutive slots on one XCD: the row tile of A (the
  // large operand) is read from HBM once and served from that L2 to all of
  // them, and only the small B operand is streamed by every XCD.
  const long long nct = npad / SPL_BN, nrt = mpad / SPL_BM;
  const long long xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  const long long rt = (slot / nct) * 8 + xcd, ct = slot % nct;
  if (rt >= nrt) return;   // whole workgroup, before any barrier
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const long long rb0 = rt * (SPL_BM / 16), cb0 = ct * (SPL_BN / 16);
  const int wr = (w & 1) * 2, wc = (w >> 1) * 4;   // the wave's first row / column block in the tile
  const long long nks = kpad / 32;
  const size_t pa = (size_t)(mpad / 16) * nks * 64, pb = (size_t)(npad / 16) * nks * 64;
  const long long row0 = rb0 * 16 + wr * 16, col0 = cb0 * 16 + wc * 16;
  double acc[2][4][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[i][j][v] = 0.0;
  stage_step(lds[0], A, B, pa, pb, rb0, cb0, nks, 0, w, lane);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (long long kc = 0; kc < nks; kc += SPL_KC / 32) {
    f32x4 l2[2][4], l3[2][4], l4[2][4], l5[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) l2[i][j] = l3[i][j] = l4[i][j] = l5[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < SPL_KC / 32; ++ks) {
      const long long kk = kc + ks;
      const int cur = (int)(kk & 1);
      if (kk + 1 < nks) stage_step(lds[cur ^ 1], A, B, pa, pb, rb0, cb0, nks, kk + 1, w, lane);
      // fragments read slice by slice and products issued level by level
      // over all eight blocks, so that each MFMA waits only for the slices
      // it uses and consecutive MFMAs update different accumulators
      bf16x8 a[2][SPL_S], b[4][SPL_S];
#pragma unroll
      for (int t = 0; t < SPL_S; ++t) {
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i][t] = lds[cur][(wr + i) * SPL_S + t][lane];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j][t] = lds[cur][(SPL_BM / 16 + wc + j) * SPL_S + t][lane];
      }
#define SPL_PRODUCT(L, T, U)                                                                          \
  _Pragma("unroll") for (int i = 0; i < 2; ++i) _Pragma("unroll") for (int j = 0; j < 4; ++j) L[i][j] = \
      __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][T], b[j][U], L[i][j], 0, 0, 0);
      SPL_PRODUCT(l2, 0, 0)
      SPL_PRODUCT(l3, 0, 1)
      SPL_PRODUCT(l3, 1, 0)
      SPL_PRODUCT(l4, 1, 1)
      SPL_PRODUCT(l4, 0, 2)
      SPL_PRODUCT(l4, 2, 0)
      SPL_PRODUCT(l5, 1, 2)
      SPL_PRODUCT(l5, 2, 1)
      SPL_PRODUCT(l5, 0, 3)
      SPL_PRODUCT(l5, 3, 0)
#undef SPL_PRODUCT
      // the next step's fragments have landed and every wave is done with
      // this step's buffer before it is refilled; the scheduling barrier
      // keeps the MFMAs above the wait (left free, the scheduler sinks them
      // below it and the wait then exposes the whole global load latency)
      __builtin_amdgcn_sched_barrier(0);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }
    const long long ci = kc / (SPL_KC / 32);
    int er[2][4], ec[4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) er[i][v] = ea[ci * mpad + row0 + 16 * i + 4 * lg + v];
#pragma unroll
    for (int j = 0; j < 4; ++j) ec[j] = eb[ci * npad + col0 + 16 * j + lr];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          acc[i][j][v] += ldexp((double)l2[i][j][v] * 0x1p-16 + (double)l3[i][j][v] * 0x1p-24 +
                                    (double)(l4[i][j][v] + l5[i][j][v] * 0x1p-8f) * 0x1p-32,
                                er[i][v] + ec[j]);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long long r = row0 + 16 * i + 4 * lg + v, c = col0 + 16 * j + lr;
        if (r < m && c < n) {
          const double val = alpha * acc[i][j][v];
          double *p = C + r + ldc * c;
          *p = beta == 0.0 ? val : beta * *p + val;
        }
      }
}

static long long round_up(long long x, long long q) { return (x + q - 1) / q * q; }

// Slice the rows of one operand (element (r, k) at X[r rs + k ks]) into the
// context's workspace `which`; *e receives the chunk exponents.
static const bf16x8 *slice_operand(Context &C, int which, const double *X, long long rs, long long ks, long long rows,
                                   long long kd, long long rows_pad, long long kpad, const int **e) {
  DBuf<unsigned short> &P = C.spl_slices[which];
  DBuf<int> &ex = C.spl_exp[which];
  P.alloc((size_t)SPL_S * (size_t)rows_pad * (size_t)kpad);
  ex.alloc((size_t)rows_pad * (size_t)(kpad / SPL_KC));
  dim3 g((unsigned)(rows_pad / SPL_T), (unsigned)(kpad / SPL_KC));
  hipLaunchKernelGGL(k_slice_rows, g, dim3(256), 0, C.stream, X, rs, ks, rows, kd, rows_pad, kpad, (bf16x8 *)P.p,
                     ex.p);
  GI_HIP(hipGetLastError());
  *e = ex.p;
  return (const bf16x8 *)P.p;
}

// C (m x n) = beta C + alpha op(A) op(B), op as in rocblas (none / transpose);
// A stored ra x ca (lda), B stored rb x cb (ldb).
void split_bf16_gemm(Context &C, rocblas_operation ta, rocblas_operation tb, long long m, long long n, long long k,
                     double alpha, const double *A, long long lda, const double *B, long long ldb, double beta,
                     double *Cd, long long ldc) {
  if (m <= 0 || n <= 0) return;
  if (k <= 0) {
    if (beta != 1.0) throw Error{63, "split product with k = 0 and beta != 1"};
    return;
  }
  const long long kpad = round_up(k, SPL_KC), mpad = round_up(m, SPL_BM), npad = round_up(n, SPL_BN);
  if (kpad / SPL_KC > 65535) throw Error{63, "split product: dimension beyond the launch grid"};
  // row r of op(A): op(A)(r, k); row c of op(B)^T: op(B)(k, c)
  const long long ars = ta == rocblas_operation_none ? 1 : lda, aks = ta == rocblas_operation_none ? lda : 1;
  const long long brs = tb == rocblas_operation_none ? ldb : 1, bks = tb == rocblas_operation_none ? 1 : ldb;
  const int *ea = nullptr, *eb = nullptr;
  const bf16x8 *pa = slice_operand(C, 0, A, ars, aks, m, k, mpad, kpad, &ea);
  const bf16x8 *pb = slice_operand(C, 1, B, brs, bks, n, k, npad, kpad, &eb);
  const long long nrt = mpad / SPL_BM, nct = npad / SPL_BN;
  const long long nwg = 8 * ((nrt + 7) / 8) * nct;
  if (nwg > 0x7fffffffLL) throw Error{63, "split product: dimension beyond the launch grid"};
  dim3 grid((unsigned)nwg);
  hipLaunchKernelGGL(k_split_gemm, grid, dim3(256), 0, C.stream, pa, pb, ea, eb, mpad, npad, kpad, m, n, alpha, beta,
                     Cd, ldc);
  GI_HIP(hipGetLastError());
}

}  // namespace gi
