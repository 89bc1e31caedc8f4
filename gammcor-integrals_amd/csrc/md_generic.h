// Hermite-matrix ERI kernel for high angular momentum: bra or ket pairs with
// l0 + l1 > 6 (g and h shells, the reference's Auto2e classes up to
// AUTO2E_MAXL = 5, auto2e.f90:275,2876-2880).  The same two products as
// md_kernel,
//   stage 1  X(h, c) = sum_k (-1)^{|k|} R_{h+k} ET_ket(k, c)
//   stage 2  V(r, c) = sum_h ET_bra(h, r) X(h, c),
// with the Hermite orders E, F as runtime arguments instead of template
// parameters (one kernel for every high class rather than 72 instantiations
// whose unrolled register tiles -- up to 286 x 441 -- would not fit).  One
// 256-lane workgroup per (bra pair, ket pair) quartet: the R table of order
// E + F (<= 20) in LDS (ping-pong recursion), X for GEN_CC ket columns at a
// time in LDS, accumulated over the ket primitive pairs, and the V tile in
// registers, accumulated over the bra primitive pairs, so contracted pairs
// cost no global read-modify-write.  FP64 VALU throughout; these classes are
// rare next to the l <= 3 work of the MFMA kernels.
#pragma once
#include "md_kernel.h"

namespace gi {

constexpr int GEN_NT = 256;                  // lanes per workgroup
constexpr int GEN_CC = 16;                   // ket columns per chunk
constexpr int GEN_EMAX = 2 * MAXL + 1;       // highest bra/ket Hermite order (11: gradient bra of an hh pair)
constexpr int GEN_NHMAX = nherm(GEN_EMAX);   // 364
constexpr int GEN_VPT = (nab_max(2 * MAXL) * GEN_CC + GEN_NT - 1) / GEN_NT;   // V elements per lane

// F_0..F_M at x into F (LDS), runtime M <= BOYS_MAXM (boys.f90:202-281, the
// same interpolation and recursions as boys_fm).  Lane 0 only.
__device__ inline void boys_fm_rt(double x, int M, const double *__restrict__ tab, double *F) {
  constexpr double dx = BOYS_XASYM / (double)BOYS_NINT;
  constexpr double inv_dx = (double)BOYS_NINT / BOYS_XASYM;
  if (x < BOYS_XASYM) {
    int k = (int)floor(x / dx);
    if (k > BOYS_NINT - 1) k = BOYS_NINT - 1;
    if (k < 0) k = 0;
    const double *c = tab + 5 * (k + BOYS_NINT * M);
    const double ab_sum = (double)(2 * k + 1) * dx;
    const double y = (2.0 * x - ab_sum) * inv_dx;
    const double twoy = 2.0 * y;
    const double d5 = c[4];
    const double d4v = twoy * d5 + c[3];
    const double d3 = twoy * d4v - d5 + c[2];
    const double d2 = twoy * d3 - d4v + c[1];
    double f = y * d2 - d3 + 0.5 * c[0];
    F[M] = f;
    const double expmx = exp(-x), twox = 2.0 * x;
    for (int l = M; l >= 1; --l) {
      f = (twox * f + expmx) * (1.0 / (double)(2 * l - 1));
      F[l - 1] = f;
    }
  } else {
    double f = 0.5 * sqrt(3.141592653589793 / x);
    F[0] = f;
    const double expmx = exp(-x), beta = 1.0 / (2.0 * x);
    for (int l = 1; l <= M; ++l) {
      f = beta * ((double)(2 * l - 1) * f - expmx);
      F[l] = f;
    }
  }
}

// R_{tuv} of one primitive quartet, order L, into the returned buffer (one of
// Ra, Rb).  Ends with a barrier.
__device__ inline const double *rtable_rt(int L, double *Ra, double *Rb, double *sF, const PrimRec &pb,
                                          const PrimRec &pk, double kappa, const double *__restrict__ boys) {
  const int tid = threadIdx.x;
  const double PQx = pb.px - pk.px, PQy = pb.py - pk.py, PQz = pb.pz - pk.pz;
  __syncthreads();   // earlier readers of Ra, Rb, sF are done
  if (tid == 0) {
    const double pq = pb.p * pk.p;
    const double den = pb.p + pk.p + kappa * pq;
    const double alpha = pq / den;
    boys_fm_rt(alpha * (PQx * PQx + PQy * PQy + PQz * PQz), L, boys, sF);
    const double m2a = -2.0 * alpha;
    double pw = 34.986836655249725 / (pq * sqrt(den));   // 2 pi^(5/2) / (p q sqrt(p + q + kappa p q))
    for (int n = 0; n <= L; ++n) {
      sF[n] *= pw;
      pw *= m2a;
    }
    ((L % 2 == 0) ? Ra : Rb)[0] = sF[L];
  }
  for (int n = L - 1; n >= 0; --n) {
    double *cur = (n % 2 == 0) ? Ra : Rb;
    const double *prv = (n % 2 == 0) ? Rb : Ra;
    __syncthreads();
    const int cnt = nherm(L - n);
    for (int idx = tid; idx < cnt; idx += GEN_NT) {
      double r;
      if (idx == 0) {
        r = sF[n];
      } else {
        const int rec = c_rrec[idx];
        const int i1 = rec & 0x7ff, i2 = (rec >> 11) & 0x7ff, dir = (rec >> 22) & 3, coef = rec >> 24;
        const double pqd = dir == 0 ? PQx : (dir == 1 ? PQy : PQz);
        r = (double)coef * prv[i2] + pqd * prv[i1];
      }
      cur[idx] = r;
    }
  }
  __syncthreads();
  return Ra;
}

__global__ void __launch_bounds__(GEN_NT) md_generic_kernel(MdArgs A, int E, int F) {
  __shared__ double s_R[2][NHRMAX];
  __shared__ double s_X[GEN_NHMAX][GEN_CC];
  __shared__ double s_F[BOYS_MAXM + 1];
  const int tid = threadIdx.x;
  const int NHE = nherm(E), NHF = nherm(F), L = E + F;
  const long long ntiles = A.ntiles;
  for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
    long long ir, jc;
    if (A.mode == LAUNCH_PRODUCT) {
      ir = t / A.ncols;
      jc = t - ir * A.ncols;
    } else {
      ir = t;
      jc = t;
    }
    const QRec RB = A.rrec[ir], CD = A.crec[jc];
    const int nr = RB.nrow, nc = CD.nrow;
    const int ldb = RB.n0 * RB.n1, ldk = CD.n0 * CD.n1;
    double *const out = A.out + RB.off;
    const bool diag = A.mode == LAUNCH_DIAG;
    for (int c0 = 0; c0 < nc; c0 += GEN_CC) {
      const int ncs = (nc - c0) < GEN_CC ? (nc - c0) : GEN_CC;
      double vacc[GEN_VPT];
#pragma unroll
      for (int i = 0; i < GEN_VPT; ++i) vacc[i] = 0.0;
      for (int bp = 0; bp < RB.nprim; ++bp) {
        const PrimRec PB = bp == 0 ? first_prim(RB) : A.prim[RB.prim0 + bp];
        const double *etb = A.et + PB.et;
        __syncthreads();   // readers of s_X from the previous stage 2 are done
        for (int e = tid; e < NHE * GEN_CC; e += GEN_NT) (&s_X[0][0])[e] = 0.0;
        for (int kp = 0; kp < CD.nprim; ++kp) {
          const PrimRec PK = kp == 0 ? first_prim(CD) : A.prim[CD.prim0 + kp];
          const double *Rf = rtable_rt(L, s_R[0], s_R[1], s_F, PB, PK, A.kappa, A.boys);
          const double *etk = A.et + PK.et;
          // stage 1: lane (h, cc) with cc fastest
          for (int e = tid; e < NHE * GEN_CC; e += GEN_NT) {
            const int h = e / GEN_CC, cc = e - h * GEN_CC;
            if (cc >= ncs) continue;
            const int fc = pair_full(CD, A.sel, c0 + cc);
            const int th = c_herm[h];
            const int t1 = px(th), u1 = py(th), v1 = pz(th);
            double acc = 0.0;
            for (int k = 0; k < NHF; ++k) {
              const int tk = c_herm[k];
              const int t = t1 + px(tk), u = u1 + py(tk), v = v1 + pz(tk);
              double r = Rf[hidx(t, u, v)];
              if ((px(tk) + py(tk) + pz(tk)) & 1) r = -r;
              acc += r * etk[(long long)k * ldk + fc];
            }
            s_X[h][cc] += acc;
          }
        }
        __syncthreads();
        // stage 2: lane (r, cc) with cc fastest; diagonal mode: r = the column
#pragma unroll
        for (int i = 0; i < GEN_VPT; ++i) {
          const int e = tid + GEN_NT * i;
          const int r = diag ? c0 + e : e / GEN_CC, cc = diag ? e : e - (e / GEN_CC) * GEN_CC;
          const bool ok = diag ? (e < ncs) : (r < nr && cc < ncs);
          if (!ok) continue;
          const int fr = pair_full(RB, A.sel, r);
          double v = 0.0;
          for (int h = 0; h < NHE; ++h) v += etb[(long long)h * ldb + fr] * s_X[h][cc];
          vacc[i] += v;
        }
      }
      // store the chunk (PRODUCT / ZIP: V(r, c0 + cc); DIAG: V(c, c); rowsum:
      // each row's sum over the chunk's columns added to out[r])
#pragma unroll
      for (int i = 0; i < GEN_VPT; ++i) {
        const int e = tid + GEN_NT * i;
        if (diag) {
          if (e < ncs) out[CD.off + c0 + e] = vacc[i];
          continue;
        }
        const int r = e / GEN_CC, cc = e - r * GEN_CC;
        if (A.rowsum) {
          // the GEN_CC lanes of a row are consecutive within one wavefront
          double s = (r < nr && cc < ncs) ? vacc[i] : 0.0;
          s += __shfl_xor(s, 1);
          s += __shfl_xor(s, 2);
          s += __shfl_xor(s, 4);
          s += __shfl_xor(s, 8);
          if (cc == 0 && r < nr) atomicAdd(out + r, s);
          continue;
        }
        if (r < nr && cc < ncs) out[CD.off + A.ldo * (c0 + cc) + r] = vacc[i];
      }
    }
  }
}

}  // namespace gi
