// One-electron integrals S, T and V_ne of a shell pair (ints1e_S / ints1e_Kinetic /
// ints1e_Coulomb, OneElectronInts.f90:242-329, 518-805, behind ints1e_gammcor_H0,
// OneElectronInts_Gammcor.f90:13-73).  Included by eri.hip, where the constant
// tables c_spher, c_herm and c_rrec live.
#pragma once
#include "gi_context.h"

namespace gi {

constexpr int ONEE_LMAX = 3;                                   // shells of the basis (l <= 3)
constexpr int ONEE_NC = ncart(ONEE_LMAX);                      // Cartesian functions per shell
constexpr int ONEE_N1D = (ONEE_LMAX + 1) * (ONEE_LMAX + 3) * (2 * ONEE_LMAX + 3);

// R_tuv(p, P - C) of a point charge (no prefactor) into Ra (level 0), Rb the
// ping-pong partner; the downward recursion of rtable with alpha = p.  Starts
// and ends with a barrier.
template <int L>
__device__ __forceinline__ void rtable_point(double *Ra, double *Rb, const double P[3], double p, double cx,
                                             double cy, double cz, const double *__restrict__ boys, int lane) {
  const double X = P[0] - cx, Y = P[1] - cy, Z = P[2] - cz;
  double Fm[L + 1];
  boys_fm<L>(p * (X * X + Y * Y + Z * Z), boys, Fm);
  const double m2a = -2.0 * p;
  double pw = 1.0;
#pragma unroll
  for (int n = 0; n <= L; ++n) {
    Fm[n] *= pw;
    pw *= m2a;
  }
  __syncthreads();   // previous readers of Ra/Rb are done
  {
    double *top = (L % 2 == 0) ? Ra : Rb;
    if (lane == 0) top[0] = Fm[L];
  }
#pragma unroll
  for (int n = L - 1; n >= 0; --n) {
    double *cur = (n % 2 == 0) ? Ra : Rb;
    const double *prv = (n % 2 == 0) ? Rb : Ra;
    __syncthreads();
    const int cnt = nherm(L - n);
    for (int idx = lane; idx < cnt; idx += 64) {
      double r;
      if (idx == 0) {
        r = Fm[n];
      } else {
        const int rec = c_rrec[idx];
        const int i1 = rec & 0x7ff, i2 = (rec >> 11) & 0x7ff, dir = (rec >> 22) & 3, coef = rec >> 24;
        const double d = dir == 0 ? X : (dir == 1 ? Y : Z);
        r = (double)coef * prv[i2] + d * prv[i1];
      }
      cur[idx] = r;
    }
  }
  __syncthreads();
}

// One 64-lane workgroup per shell pair (s0 >= s1, l0 + l1 = L): Cartesian
// blocks of S, T (Ferman & Valeev eq. 4.8, ket exponent) and V (Helgaker
// 9.9.32, the R tables of all charges summed first) over the primitive
// pairs, then the AO block nA nB U0 X U1^T written to both triangles.
template <int L>
__global__ void __launch_bounds__(64) k_one_electron(const DevShell *__restrict__ shells,
                                                     const double *__restrict__ exps,
                                                     const double *__restrict__ cntr,
                                                     const double *__restrict__ norms,
                                                     const int *__restrict__ ao_loc, int spher,
                                                     const int *__restrict__ pairs,
                                                     const double *__restrict__ chg, int ncharge,
                                                     const double *__restrict__ boys, double *S, double *T,
                                                     double *V, long long ldh) {
  constexpr int NHL = nherm(L);
  __shared__ double s_e[3][ONEE_N1D];
  __shared__ double s_R[2][NHL];
  __shared__ double s_RR[NHL];
  __shared__ double s_acc[3][ONEE_NC * ONEE_NC];
  __shared__ double s_tmp[(2 * ONEE_LMAX + 1) * ONEE_NC];
  const int lane = threadIdx.x;
  const int sa = pairs[2 * blockIdx.x], sb = pairs[2 * blockIdx.x + 1];
  const DevShell A = shells[sa], B = shells[sb];
  const int la = A.l, lb = B.l;
  const int nca = ncart(la), ncb = ncart(lb);
  const int s_j = L + 3, s_i = (lb + 3) * s_j;   // E(i, j, t), j <= lb + 2, t <= L + 2
  for (int e = lane; e < 3 * ONEE_NC * ONEE_NC; e += 64) (&s_acc[0][0])[e] = 0.0;
  const double AB[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
  const double AB2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  const double Ac[3] = {A.x, A.y, A.z}, Bc[3] = {B.x, B.y, B.z};
  for (int ib = 0; ib < B.np; ++ib)
    for (int ia = 0; ia < A.np; ++ia) {
      const double a = exps[A.prim0 + ia], b = exps[B.prim0 + ib];
      const double p = a + b, mu = a * b / p;
      const double P[3] = {(a * A.x + b * B.x) / p, (a * A.y + b * B.y) / p, (a * A.z + b * B.z) / p};
      const double pref = cntr[A.prim0 + ia] * cntr[B.prim0 + ib] * exp(-mu * AB2);
      __syncthreads();   // the previous primitive pair's readers are done
      if (lane < 3) {
        double *E = s_e[lane];
        const double xpa = P[lane] - Ac[lane], xpb = P[lane] - Bc[lane], i2p = 0.5 / p;
        for (int k = 0; k < (la + 1) * s_i; ++k) E[k] = 0.0;
        E[0] = 1.0;
        for (int i = 0; i < la; ++i)
          for (int t = 0; t <= i + 1; ++t) {
            double v = t <= i ? xpa * E[i * s_i + t] : 0.0;
            if (t >= 1) v += i2p * E[i * s_i + t - 1];
            if (t + 1 <= i) v += (double)(t + 1) * E[i * s_i + t + 1];
            E[(i + 1) * s_i + t] = v;
          }
        for (int i = 0; i <= la; ++i)
          for (int j = 0; j < lb + 2; ++j)
            for (int t = 0; t <= i + j + 1; ++t) {
              const int o = i * s_i + j * s_j;
              double v = t <= i + j ? xpb * E[o + t] : 0.0;
              if (t >= 1) v += i2p * E[o + t - 1];
              if (t + 1 <= i + j) v += (double)(t + 1) * E[o + t + 1];
              E[o + s_j + t] = v;
            }
      }
      for (int h = lane; h < NHL; h += 64) s_RR[h] = 0.0;
      if (V)
        for (int c = 0; c < ncharge; ++c) {
          const double q = chg[4 * c + 3];
          rtable_point<L>(s_R[0], s_R[1], P, p, chg[4 * c], chg[4 * c + 1], chg[4 * c + 2], boys, lane);
          for (int h = lane; h < NHL; h += 64) s_RR[h] -= q * s_R[0][h];
        }
      __syncthreads();
      const double sp = pref * pow(3.141592653589793 / p, 1.5), vp = pref * 6.283185307179586 / p;
      for (int e = lane; e < nca * ncb; e += 64) {
        const int ca = e % nca, cb = e / nca;
        const int pa_ = cart_of(la, ca), pb_ = cart_of(lb, cb);
        const int ax = px(pa_), ay = py(pa_), az = pz(pa_), bx = px(pb_), by = py(pb_), bz = pz(pb_);
        const double *Ex = s_e[0], *Ey = s_e[1], *Ez = s_e[2];
        const double x0 = Ex[ax * s_i + bx * s_j], y0 = Ey[ay * s_i + by * s_j], z0 = Ez[az * s_i + bz * s_j];
        const double sab = x0 * y0 * z0;
        double Ix = (2 * bx + 1) * b * sab - 2.0 * b * b * Ex[ax * s_i + (bx + 2) * s_j] * y0 * z0;
        if (bx >= 2) Ix -= bx * (bx - 1) * 0.5 * Ex[ax * s_i + (bx - 2) * s_j] * y0 * z0;
        double Iy = (2 * by + 1) * b * sab - 2.0 * b * b * Ey[ay * s_i + (by + 2) * s_j] * x0 * z0;
        if (by >= 2) Iy -= by * (by - 1) * 0.5 * Ey[ay * s_i + (by - 2) * s_j] * x0 * z0;
        double Iz = (2 * bz + 1) * b * sab - 2.0 * b * b * Ez[az * s_i + (bz + 2) * s_j] * x0 * y0;
        if (bz >= 2) Iz -= bz * (bz - 1) * 0.5 * Ez[az * s_i + (bz - 2) * s_j] * x0 * y0;
        double v = 0.0;
        if (V)
          for (int t = 0; t <= ax + bx; ++t)
            for (int u = 0; u <= ay + by; ++u)
              for (int w = 0; w <= az + bz; ++w)
                v += Ex[ax * s_i + bx * s_j + t] * Ey[ay * s_i + by * s_j + u] * Ez[az * s_i + bz * s_j + w] *
                     s_RR[hidx(t, u, w)];
        s_acc[0][e] += sp * sab;
        s_acc[1][e] += sp * (Ix + Iy + Iz);
        s_acc[2][e] += vp * v;
      }
    }
  __syncthreads();
  // the Cartesian diagonal block takes its lower triangle (real_smfill)
  if (sa == sb) {
    for (int e = lane; e < nca * ncb; e += 64) {
      const int ca = e % nca, cb = e / nca;
      if (ca < cb)
        for (int k = 0; k < 3; ++k) s_acc[k][e] = s_acc[k][cb + nca * ca];
    }
    __syncthreads();
  }
  const bool t0 = spher && la >= 2, t1 = spher && lb >= 2;
  const int n0 = t0 ? 2 * la + 1 : nca, n1 = t1 ? 2 * lb + 1 : ncb;
  const double *U0 = c_spher + spher_off(la >= 2 ? la : 2), *U1 = c_spher + spher_off(lb >= 2 ? lb : 2);
  const double *nA = norms + A.norm0, *nB = norms + B.norm0;
  const long long p0 = ao_loc[sa], q0 = ao_loc[sb];
  double *outs[3] = {S, T, V};
  for (int k = 0; k < 3; ++k) {
    double *out = outs[k];
    if (!out) continue;   // uniform
    for (int e = lane; e < n0 * ncb; e += 64) {
      const int ma = e % n0, cb = e / n0;
      double acc = 0.0;
      for (int ca = 0; ca < nca; ++ca) {
        const double u = t0 ? U0[ma * nca + ca] : (ma == ca ? 1.0 : 0.0);
        if (u != 0.0) acc += u * s_acc[k][ca + nca * cb];
      }
      s_tmp[ma + n0 * cb] = acc;
    }
    __syncthreads();
    for (int e = lane; e < n0 * n1; e += 64) {
      const int ma = e % n0, mb = e / n0;
      if (sa == sb && ma < mb) continue;
      double acc = 0.0;
      for (int cb = 0; cb < ncb; ++cb) {
        const double u = t1 ? U1[mb * ncb + cb] : (mb == cb ? 1.0 : 0.0);
        if (u != 0.0) acc += u * s_tmp[ma + n0 * cb];
      }
      const double val = nA[ma] * nB[mb] * acc;
      out[(p0 + ma) + ldh * (q0 + mb)] = val;
      out[(q0 + mb) + ldh * (p0 + ma)] = val;
    }
    __syncthreads();
  }
}

void launch_one_electron(const Context &C, int L, const int *ao_loc, const int *pairs, long long npairs,
                         const double *chg, int ncharge, double *S, double *T, double *V, long long ldh) {
  if (npairs <= 0) return;
  const dim3 g((unsigned)npairs), b(64);
#define GI_ONEE(LL)                                                                                     \
  case LL:                                                                                              \
    hipLaunchKernelGGL(k_one_electron<LL>, g, b, 0, C.stream, C.dshells.p, C.dexps.p, C.dcntr.p, C.dnorms.p, \
                       ao_loc, C.spher, pairs, chg, ncharge, C.dboys.p, S, T, V, ldh);                \
    break;
  switch (L) {
    GI_ONEE(0)
    GI_ONEE(1)
    GI_ONEE(2)
    GI_ONEE(3)
    GI_ONEE(4)
    GI_ONEE(5)
    GI_ONEE(6)
    default:
      throw Error{64, "one-electron integrals need l <= 3"};
  }
#undef GI_ONEE
  GI_HIP(hipGetLastError());
}

}  // namespace gi
