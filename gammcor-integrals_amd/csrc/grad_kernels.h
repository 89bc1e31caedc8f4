// Density-contracted Hermite tables for the Coulomb energy gradient (the
// layout and the derivation are described in grad.hip).  Included by eri.hip,
// where the constant tables c_spher and c_herm live.
#pragma once
#include "gi_context.h"

namespace gi {

// One 64-lane workgroup per primitive pair: E^{ij}_t and dE^{ij}_t/dX_AB per
// direction (the two-centre recursion of k_build_et and its derivative with
// X_PA = -(b/p) X_AB, X_PB = (a/p) X_AB), the density block moved to
// Cartesian components with the AO normalization and the solid-harmonic
// transform, then the GRAD_NCOL columns above.
__global__ void __launch_bounds__(64) k_build_grad_et(const DevShell *__restrict__ shells,
                                                      const double *__restrict__ exps,
                                                      const double *__restrict__ cntr,
                                                      const double *__restrict__ norms,
                                                      const int *__restrict__ ao_loc, int spher,
                                                      const double *__restrict__ D, long long ldd,
                                                      const GradJob *__restrict__ jobs, PrimRec *prims,
                                                      double *et) {
  constexpr int N1D = (MAXL + 1) * (MAXL + 1) * (2 * MAXL + 1);
  constexpr int NC = ncart(MAXL);
  __shared__ double s_e1[3][N1D], s_d1[3][N1D];
  __shared__ double s_T[NC * NC], s_Dc[NC * NC];
  __shared__ double s_D0[nherm(2 * MAXL + 1)];
  const int lane = threadIdx.x;
  const GradJob J = jobs[blockIdx.x];
  const DevShell A = shells[J.s0], B = shells[J.s1];
  const int la = A.l, lb = B.l, lab = la + lb;
  const double a = exps[A.prim0 + J.ia], b = exps[B.prim0 + J.ib];
  const double p = a + b, mu = a * b / p;
  const double XAB[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
  const double AB2 = XAB[0] * XAB[0] + XAB[1] * XAB[1] + XAB[2] * XAB[2];
  const double pref = cntr[A.prim0 + J.ia] * cntr[B.prim0 + J.ib] * exp(-mu * AB2);
  if (lane == 0)
    prims[J.prim] = PrimRec{(a * A.x + b * B.x) / p, (a * A.y + b * B.y) / p, (a * A.z + b * B.z) / p, p, J.et, 0};
  const int s_i = (lb + 1) * (lab + 1), s_j = lab + 1;   // E1(i, j, t) at i*s_i + j*s_j + t
  if (lane < 3) {
    const double xab = XAB[lane];
    const double xpa = -(b / p) * xab, xpb = (a / p) * xab, i2p = 0.5 / p;
    const double dxpa = -b / p, dxpb = a / p;
    double *E1 = s_e1[lane], *D1 = s_d1[lane];
    for (int k = 0; k < (la + 1) * s_i; ++k) {
      E1[k] = 0.0;
      D1[k] = 0.0;
    }
    E1[0] = 1.0;
    for (int i = 0; i < la; ++i)
      for (int t = 0; t <= i + 1; ++t) {
        const double e = t <= i ? E1[i * s_i + t] : 0.0, d = t <= i ? D1[i * s_i + t] : 0.0;
        double v = xpa * e, dv = xpa * d + dxpa * e;
        if (t >= 1) {
          v += i2p * E1[i * s_i + t - 1];
          dv += i2p * D1[i * s_i + t - 1];
        }
        if (t + 1 <= i) {
          v += (double)(t + 1) * E1[i * s_i + t + 1];
          dv += (double)(t + 1) * D1[i * s_i + t + 1];
        }
        E1[(i + 1) * s_i + t] = v;
        D1[(i + 1) * s_i + t] = dv;
      }
    for (int i = 0; i <= la; ++i)
      for (int j = 0; j < lb; ++j)
        for (int t = 0; t <= i + j + 1; ++t) {
          const int o = i * s_i + j * s_j;
          const double e = t <= i + j ? E1[o + t] : 0.0, d = t <= i + j ? D1[o + t] : 0.0;
          double v = xpb * e, dv = xpb * d + dxpb * e;
          if (t >= 1) {
            v += i2p * E1[o + t - 1];
            dv += i2p * D1[o + t - 1];
          }
          if (t + 1 <= i + j) {
            v += (double)(t + 1) * E1[o + t + 1];
            dv += (double)(t + 1) * D1[o + t + 1];
          }
          E1[o + s_j + t] = v;
          D1[o + s_j + t] = dv;
        }
  }
  // density block in Cartesian components: Dc(ca, cb) = w sum_{ma,mb}
  // U0(ma, ca) nA(ma) D(p0 + ma, q0 + mb) nB(mb) U1(mb, cb)
  const int nc0 = ncart(la), nc1 = ncart(lb);
  const bool t0 = spher && la >= 2, t1 = spher && lb >= 2;
  const int n0 = t0 ? 2 * la + 1 : nc0, n1 = t1 ? 2 * lb + 1 : nc1;
  const double *U0 = c_spher + spher_off(la >= 2 ? la : 2);
  const double *U1 = c_spher + spher_off(lb >= 2 ? lb : 2);
  const double *nA = norms + A.norm0, *nB = norms + B.norm0;
  const long long p0 = ao_loc[J.s0], q0 = ao_loc[J.s1];
  for (int e = lane; e < n0 * nc1; e += 64) {
    const int ma = e % n0, cb = e / n0;
    double acc = 0.0;
    for (int mb = 0; mb < n1; ++mb) {
      const double u = t1 ? U1[mb * nc1 + cb] : (mb == cb ? 1.0 : 0.0);
      if (u != 0.0) acc += D[(p0 + ma) + ldd * (q0 + mb)] * nB[mb] * u;
    }
    s_T[ma + n0 * cb] = acc;
  }
  __syncthreads();
  for (int e = lane; e < nc0 * nc1; e += 64) {
    const int ca = e % nc0, cb = e / nc0;
    double acc = 0.0;
    for (int ma = 0; ma < n0; ++ma) {
      const double u = t0 ? U0[ma * nc0 + ca] : (ma == ca ? 1.0 : 0.0);
      if (u != 0.0) acc += u * nA[ma] * s_T[ma + n0 * cb];
    }
    s_Dc[ca + nc0 * cb] = J.w * acc;
  }
  __syncthreads();
  const int NH1 = nherm(lab + 1), NH0 = nherm(lab);
  double *out = et + J.et;
  for (int h = lane; h < NH1; h += 64) {
    const int th = c_herm[h];
    const int tt = px(th), uu = py(th), vv = pz(th);
    double d0 = 0.0, dx = 0.0, dy = 0.0, dz = 0.0;
    if (h < NH0) {
      for (int ca = 0; ca < nc0; ++ca) {
        const int cca = cart_of(la, ca);
        const int ax = px(cca), ay = py(cca), az = pz(cca);
        for (int cb = 0; cb < nc1; ++cb) {
          const int ccb = cart_of(lb, cb);
          const int bx = px(ccb), by = py(ccb), bz = pz(ccb);
          if (tt > ax + bx || uu > ay + by || vv > az + bz) continue;
          const double dc = s_Dc[ca + nc0 * cb];
          const int ix = ax * s_i + bx * s_j + tt, iy = ay * s_i + by * s_j + uu, iz = az * s_i + bz * s_j + vv;
          const double ex = s_e1[0][ix], ey = s_e1[1][iy], ez = s_e1[2][iz];
          d0 += dc * ex * ey * ez;
          dx += dc * s_d1[0][ix] * ey * ez;
          dy += dc * ex * s_d1[1][iy] * ez;
          dz += dc * ex * ey * s_d1[2][iz];
        }
      }
      d0 *= pref;
      dx = pref * dx - 2.0 * mu * XAB[0] * d0;
      dy = pref * dy - 2.0 * mu * XAB[1] * d0;
      dz = pref * dz - 2.0 * mu * XAB[2] * d0;
    }
    s_D0[h] = d0;
    out[(long long)h * GRAD_NCOL + 0] = d0;
    out[(long long)h * GRAD_NCOL + 1] = dx;
    out[(long long)h * GRAD_NCOL + 2] = dy;
    out[(long long)h * GRAD_NCOL + 3] = dz;
  }
  __syncthreads();
  const double ap = a / p;
  for (int h = lane; h < NH1; h += 64) {
    const int th = c_herm[h];
    const int tt = px(th), uu = py(th), vv = pz(th);
    const double sx = tt > 0 ? s_D0[hidx(tt - 1, uu, vv)] : 0.0;
    const double sy = uu > 0 ? s_D0[hidx(tt, uu - 1, vv)] : 0.0;
    const double sz = vv > 0 ? s_D0[hidx(tt, uu, vv - 1)] : 0.0;
    double *o = out + (long long)h * GRAD_NCOL;
    o[4] = sx;
    o[5] = sy;
    o[6] = sz;
    o[7] = ap * sx;
    o[8] = ap * sy;
    o[9] = ap * sz;
  }
}

void launch_build_grad_et(const Context &C, const int *ao_loc, const double *D, long long ldd,
                          const GradJob *jobs, long long njobs, PrimRec *prims, double *et) {
  hipLaunchKernelGGL(k_build_grad_et, dim3((unsigned)njobs), dim3(64), 0, C.stream, C.dshells.p, C.dexps.p,
                     C.dcntr.p, C.dnorms.p, ao_loc, C.spher, D, ldd, jobs, prims, et);
  GI_HIP(hipGetLastError());
}

}  // namespace gi
