// Split-BF16 products on MFMA: the precision form of the library's FP64
// GEMMs (the Cholesky trailing update and the AO->MO transform).
//
// Every FP64 operand element is split into three BF16 parts,
//   x = x1 + x2 + x3 + O(2^-24 |x|),  x1 = bf16(x), x2 = bf16(x - x1), x3 = bf16(x - x1 - x2),
// and a product is assembled from the six part products whose size is above
// 2^-24 of the leading one,
//   A B = A1 B1 + (A1 B2 + A2 B1 + A1 B3 + A2 B2 + A3 B1) + O(2^-24 |A||B|),
// each a rocBLAS gemm_ex with BF16 inputs and FP32 accumulation.  The five
// correction products go into their own FP32 accumulator, so that their
// rounding is not absorbed by the leading term, and both FP32 partials are
// added to the FP64 destination (compensated sum).  The error of the result
// is that of FP32 accumulation over k terms of |A||B|, which the callers
// report against their FP64 path.
#include <hip/hip_bf16.h>
#include "gi_context.h"

namespace gi {

#define SPLIT_RB(x)                                                                                \
  do {                                                                                           \
    rocblas_status s_ = (x);                                                                     \
    if (s_ != rocblas_status_success)                                                            \
      throw Error{2, std::string("rocBLAS error ") + rocblas_status_to_string(s_) + " at " + __FILE__ + ":" + \
                         std::to_string(__LINE__)};                                              \
  } while (0)

// parts of an m x n column-major block (leading dimension ld) into three
// contiguous m x n BF16 arrays
__global__ void k_split3(const double *__restrict__ x, long long ld, long long m, long long n,
                         __hip_bfloat16 *__restrict__ p1, __hip_bfloat16 *__restrict__ p2,
                         __hip_bfloat16 *__restrict__ p3) {
  const long long total = m * n;
  for (long long e = threadIdx.x + blockIdx.x * (long long)blockDim.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long i = e % m, j = e / m;
    const double v = x[i + ld * j];
    const __hip_bfloat16 a = __float2bfloat16((float)v);
    const double r1 = v - (double)__bfloat162float(a);
    const __hip_bfloat16 b = __float2bfloat16((float)r1);
    const double r2 = r1 - (double)__bfloat162float(b);
    p1[e] = a;
    p2[e] = b;
    p3[e] = __float2bfloat16((float)r2);
  }
}

// C = beta C + alpha (P + Q), C FP64 (ldc), P and Q FP32 contiguous m x n
__global__ void k_add_partials(double *__restrict__ C, long long ldc, long long m, long long n, double alpha,
                               double beta, const float *__restrict__ P, const float *__restrict__ Q) {
  const long long total = m * n;
  for (long long e = threadIdx.x + blockIdx.x * (long long)blockDim.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long i = e % m, j = e / m;
    const double s = (double)P[e] + (double)Q[e];
    double *c = C + i + ldc * j;
    *c = beta == 0.0 ? alpha * s : beta * *c + alpha * s;
  }
}

static unsigned grid_for(long long total) {
  return (unsigned)std::min<long long>((total + 255) / 256, 65536);
}

// C (m x n) = beta C + alpha op(A) op(B), op as in rocblas (none / transpose);
// A stored ra x ca (lda), B stored rb x cb (ldb).
void split_bf16_gemm(Context &C, rocblas_operation ta, rocblas_operation tb, long long m, long long n, long long k,
                     double alpha, const double *A, long long lda, const double *B, long long ldb, double beta,
                     double *Cd, long long ldc) {
  if (m <= 0 || n <= 0) return;
  const long long ra = ta == rocblas_operation_none ? m : k, ca = ta == rocblas_operation_none ? k : m;
  const long long rb = tb == rocblas_operation_none ? k : n, cb = tb == rocblas_operation_none ? n : k;
  if (k <= 0) {
    if (beta != 1.0) throw Error{63, "split product with k = 0 and beta != 1"};
    return;
  }
  DBuf<__hip_bfloat16> a1, a2, a3, b1, b2, b3;
  DBuf<float> P, Q;
  a1.alloc((size_t)(ra * ca));
  a2.alloc((size_t)(ra * ca));
  a3.alloc((size_t)(ra * ca));
  b1.alloc((size_t)(rb * cb));
  b2.alloc((size_t)(rb * cb));
  b3.alloc((size_t)(rb * cb));
  P.alloc((size_t)(m * n));
  Q.alloc((size_t)(m * n));
  hipLaunchKernelGGL(k_split3, dim3(grid_for(ra * ca)), dim3(256), 0, C.stream, A, lda, ra, ca, a1.p, a2.p, a3.p);
  hipLaunchKernelGGL(k_split3, dim3(grid_for(rb * cb)), dim3(256), 0, C.stream, B, ldb, rb, cb, b1.p, b2.p, b3.p);
  GI_HIP(hipGetLastError());
  const float one = 1.0f, zero = 0.0f;
  auto prod = [&](const __hip_bfloat16 *x, const __hip_bfloat16 *y, float *out, const float *bt) {
    SPLIT_RB(rocblas_gemm_ex(C.blas, ta, tb, (rocblas_int)m, (rocblas_int)n, (rocblas_int)k, &one, x,
                             rocblas_datatype_bf16_r, (rocblas_int)ra, y, rocblas_datatype_bf16_r, (rocblas_int)rb, bt,
                             out, rocblas_datatype_f32_r, (rocblas_int)m, out, rocblas_datatype_f32_r, (rocblas_int)m,
                             rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0));
  };
  // corrections, smallest first, then the leading product
  prod(a3.p, b1.p, Q.p, &zero);
  prod(a2.p, b2.p, Q.p, &one);
  prod(a1.p, b3.p, Q.p, &one);
  prod(a2.p, b1.p, Q.p, &one);
  prod(a1.p, b2.p, Q.p, &one);
  prod(a1.p, b1.p, P.p, &zero);
  hipLaunchKernelGGL(k_add_partials, dim3(grid_for(m * n)), dim3(256), 0, C.stream, Cd, ldc, m, n, alpha, beta, P.p,
                     Q.p);
  GI_HIP(hipGetLastError());
}

}  // namespace gi
