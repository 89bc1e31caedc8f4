// Atomic orbitals on grid points (gridfunc_Orbitals / gridfunc_Orbitals_Shell,
// GridFunctions.f90:154-240, 313-390): for each shell the contracted radial
// part sum_i c_i exp(-alpha_i |r - A|^2), the Cartesian monomials, the
// spherical transform for l >= 2 and the normalization factors.  Included by
// eri.hip, where the constant table c_spher lives.
#pragma once
#include "gi_context.h"

namespace gi {

// The 2l+1 (spherical, l >= 2) or ncart(l) values of one shell at one point,
// written to out0[(loc + m) ldx]
template <int L>
__device__ __forceinline__ void orbitals_of_shell(const DevShell &A, const double *__restrict__ exps,
                                                  const double *__restrict__ cntr, const double *__restrict__ norms,
                                                  int loc, int spher, double gx, double gy, double gz,
                                                  double *__restrict__ out0, long long ldx) {
  constexpr int NC = ncart(L);
  const double dx = gx - A.x, dy = gy - A.y, dz = gz - A.z;
  const double r2 = dx * dx + dy * dy + dz * dz;
  double rad = 0.0;
  for (int i = 0; i < A.np; ++i) rad += cntr[A.prim0 + i] * exp(-exps[A.prim0 + i] * r2);
  double pwx[L + 1], pwy[L + 1], pwz[L + 1];
  pwx[0] = pwy[0] = pwz[0] = 1.0;
#pragma unroll
  for (int k = 1; k <= L; ++k) {
    pwx[k] = pwx[k - 1] * dx;
    pwy[k] = pwy[k - 1] * dy;
    pwz[k] = pwz[k - 1] * dz;
  }
  double cart[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int e = cart_of(L, c);
    cart[c] = pwx[px(e)] * pwy[py(e)] * pwz[pz(e)] * rad;
  }
  const double *nA = norms + A.norm0;
  double *out = out0 + (long long)loc * ldx;
  if (L >= 2 && spher) {
    const double *U = c_spher + spher_off(L >= 2 ? L : 2);
#pragma unroll
    for (int m = 0; m < 2 * L + 1; ++m) {
      double v = 0.0;
#pragma unroll
      for (int c = 0; c < NC; ++c) v += U[m * NC + c] * cart[c];
      out[m * ldx] = nA[m] * v;
    }
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c * ldx] = nA[c] * cart[c];
  }
}

// One thread per grid point; the ns shells of angular momentum L listed in
// shl are taken blockIdx.y, blockIdx.y + gridDim.y, ...  X(g, p) column-major
// with leading dimension ldx (AO p = ao_loc[shell] + m); xyz (3, npts).
template <int L>
__global__ void __launch_bounds__(256) k_thc_orbitals(const DevShell *__restrict__ shells,
                                                      const double *__restrict__ exps,
                                                      const double *__restrict__ cntr,
                                                      const double *__restrict__ norms,
                                                      const int *__restrict__ ao_loc, const int *__restrict__ shl,
                                                      long long ns, int spher, const double *__restrict__ xyz,
                                                      long long npts, double *__restrict__ X, long long ldx) {
  const long long g = blockIdx.x * 256LL + threadIdx.x;
  if (g >= npts) return;
  const double gx = xyz[3 * g], gy = xyz[3 * g + 1], gz = xyz[3 * g + 2];
  for (long long k = blockIdx.y; k < ns; k += gridDim.y) orbitals_of_shell<L>(shells[shl[k]], exps, cntr, norms,
                                                                             ao_loc[shl[k]], spher, gx, gy, gz,
                                                                             X + g, ldx);
}

void launch_thc_orbitals(const Context &C, int L, const int *ao_loc, const int *shl, long long ns,
                         const double *xyz, long long npts, double *X, long long ldx) {
  if (ns <= 0 || npts <= 0) return;
  const dim3 g((unsigned)((npts + 255) / 256), (unsigned)std::min<long long>(ns, 65535)), b(256);
#define GI_THC_ORB(LL)                                                                                        \
  case LL:                                                                                                    \
    hipLaunchKernelGGL(k_thc_orbitals<LL>, g, b, 0, C.stream, C.dshells.p, C.dexps.p, C.dcntr.p, C.dnorms.p, \
                       ao_loc, shl, ns, C.spher, xyz, npts, X, ldx);                                          \
    break;
  switch (L) {
    GI_THC_ORB(0)
    GI_THC_ORB(1)
    GI_THC_ORB(2)
    GI_THC_ORB(3)
    GI_THC_ORB(4)
    GI_THC_ORB(5)
    default:
      throw Error{80, "orbitals on grid points need l <= 5"};
  }
#undef GI_THC_ORB
  GI_HIP(hipGetLastError());
}

}  // namespace gi
