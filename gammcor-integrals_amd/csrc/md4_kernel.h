// Hermite-matrix ERI kernel on the 4x4x4 FP64 matrix instruction
// (v_mfma_f64_4x4x4_4b_f64, four independent 4x4x4 blocks per instruction).
//
// Same two products as md_kernel.h,
//   stage 1  X(h, c) = sum_k (-1)^{|k|} R_{h+k} ET_ket(k, c)
//   stage 2  V(r, c) = sum_h ET_bra(h, r) X(h, c),
// (Helgaker, Joergensen, Olsen eqs. 9.9.18-20), organised for thin ket sides:
// one wavefront takes one bra pair and a column tile of four slots, each slot
// four ket columns of one ket pair; block b of every instruction serves
// slot b.  The 4x4x4 instruction's lane layout on gfx950 (measured): lane
// l = 16 q + 4 b + i holds A_b(i, q), B_b(q, i) and D_b(q, i), so the stage-1
// result D = X(h, c) is already the stage-2 B operand and stays in registers.
// Padding is to multiples of four (h, k, rows, columns) instead of sixteen.
//
// R_{tuv} of each quartet is stored densely, R[(t (L+1) + u) (L+1) + v], so
// that the address of R_{h+k} is off(h) + off(k): every stage-1 operand is
// one LDS read at a per-lane offset, no index table.  The R tables of the
// tile's distinct quartets (up to four, fewer for high orders to bound LDS)
// are built together, all lanes over (quartet, entry) per recursion level.
// Contracted pairs: X accumulates over ket primitives in registers; the
// block of a contracted bra accumulates over its primitive pairs in place.
#pragma once
#include "md_kernel.h"

namespace gi {

// column tile of a launch: slot s takes columns c0[s] .. c0[s] + 3 of ket
// pair jq[rt[s]] (positions in the launch's ket list); c0 = -1: empty slot
struct ColTile {
  int jq[4];
  int c0[4];
  int rt[4];
  int nq;          // distinct ket pairs jq[0 .. nq-1]
  int pad[3];
};

__host__ __device__ constexpr int m4_dsize(int L) { return (L + 1) * (L + 1) * (L + 1); }
// distinct quartets per tile: their dense R tables within ~12 KB of LDS
__host__ __device__ constexpr int m4_qmax(int L) {
  return imax(1, imin(4, 12288 / (8 * m4_dsize(L))));
}

#ifndef MD4_ATTR
#define MD4_ATTR
#endif

__device__ __forceinline__ int pick4(const int (&v)[4], int s) {
  return s == 0 ? v[0] : (s == 1 ? v[1] : (s == 2 ? v[2] : v[3]));
}

template <int E, int F>
__global__ void __launch_bounds__(64) MD4_ATTR md4_kernel(MdArgs A, const ColTile *__restrict__ ctiles, long long nct) {
  constexpr int L = E + F, L1 = L + 1;
  constexpr int NHE = nherm(E), NHF = nherm(F), NHL = nherm(L);
  constexpr int NHG = (NHE + 3) / 4, KSF = (NHF + 3) / 4;
  constexpr int DS = m4_dsize(L), QM = m4_qmax(L), NL1 = L >= 1 ? nherm(L - 1) : 1;
  __shared__ double s_Rd[QM][DS];        // level-0 R tables, dense
  __shared__ double s_Rp[2][QM][NL1];    // levels >= 1, packed (ping-pong)
  __shared__ double s_F[QM][L + 1];      // scaled Boys values per quartet
  __shared__ double s_PQ[QM][4];         // P - Q, active flag
  __shared__ int s_rrec[NHL];
  __shared__ unsigned short s_dix[NHL];  // dense offset of packed index
  __shared__ int s_hoff[NHG * 4];        // dense offset of h (0 for padding)

  const int lane = threadIdx.x;
  const int lq = lane >> 4, lb = (lane >> 2) & 3, li = lane & 3;
  for (int e = lane; e < NHL; e += 64) {
    s_rrec[e] = c_rrec[e];
    const int th = c_herm[e];
    s_dix[e] = (unsigned short)((px(th) * L1 + py(th)) * L1 + pz(th));
  }
  for (int e = lane; e < NHG * 4; e += 64) {
    int o = 0;
    if (e < NHE) {
      const int th = c_herm[e];
      o = (px(th) * L1 + py(th)) * L1 + pz(th);
    }
    s_hoff[e] = o;
  }
  for (int e = lane; e < QM * DS; e += 64) (&s_Rd[0][0])[e] = 0.0;
  for (int e = lane; e < 2 * QM * NL1; e += 64) (&s_Rp[0][0][0])[e] = 0.0;
  // stage-1 operand offsets of this lane's k = 4 kg + lq and their signs
  int koff[KSF];
  unsigned ksgn = 0;
#pragma unroll
  for (int kg = 0; kg < KSF; ++kg) {
    const int k = 4 * kg + lq;
    koff[kg] = 0;
    if (k < NHF) {
      const int tk = c_herm[k];
      koff[kg] = (px(tk) * L1 + py(tk)) * L1 + pz(tk);
      if ((px(tk) + py(tk) + pz(tk)) & 1) ksgn |= 1u << kg;
    }
  }
  __syncthreads();
  int hoff[NHG];   // dense offset of this lane's h = 4 hg + li
#pragma unroll
  for (int hg = 0; hg < NHG; ++hg) hoff[hg] = s_hoff[4 * hg + li];

  // XCD-aware order: consecutive tiles (one bra pair) on one XCD's L2
  const int G = gridDim.x;
  const int w = blockIdx.x;
  const long long v0 = (G % 8 == 0) ? (long long)(w % 8) * (G / 8) + w / 8 : w;
  const long long rs = A.trans ? A.ldo : 1, cs = A.trans ? 1 : A.ldo;

  for (long long t = v0; t < A.ntiles; t += G) {
    const long long ir = t / nct, ct = t - ir * nct;
    const QRec RB = A.rrec[ir];
    const ColTile T = ctiles[ct];
    const int nq = T.nq;
    // this lane's slot: ket pair, column
    const int myrt = pick4(T.rt, lb);
    const int myc0 = pick4(T.c0, lb);
    const QRec CD = A.crec[pick4(T.jq, myrt)];
    const int c = myc0 + li;
    const bool cval = myc0 >= 0 && c < CD.nrow;
    const int fc = cval ? pair_full(CD, A.sel, c) : 0;
    const int ldk = CD.n0 * CD.n1;
    int maxnp = 1;
    for (int qq = 0; qq < nq; ++qq) maxnp = imax(maxnp, A.crec[pick4(T.jq, qq)].nprim);
    const int nr = RB.nrow, ldb = RB.n0 * RB.n1;
    double *const out = A.out + RB.off + CD.off;
    const int nrg = (nr + 3) / 4;

    for (int bp = 0; bp < RB.nprim; ++bp) {
      const PrimRec PB = bp == 0 ? first_prim(RB) : A.prim[RB.prim0 + bp];
      const double *etb = A.et + PB.et;
      double X[NHG];
#pragma unroll
      for (int hg = 0; hg < NHG; ++hg) X[hg] = 0.0;

      for (int kp = 0; kp < maxnp; ++kp) {
        __syncthreads();   // earlier readers of the R tables are done
        // Boys function and top level of each distinct quartet
        if (lane < nq) {
          const QRec CQ = A.crec[pick4(T.jq, lane)];
          const bool act = kp < CQ.nprim;
          double PQx = 0.0, PQy = 0.0, PQz = 0.0;
          if (act) {
            const PrimRec PK = kp == 0 ? first_prim(CQ) : A.prim[CQ.prim0 + kp];
            PQx = PB.px - PK.px;
            PQy = PB.py - PK.py;
            PQz = PB.pz - PK.pz;
            const double pq = PB.p * PK.p;
            const double den = PB.p + PK.p + A.kappa * pq;
            const double alpha = pq / den;
            const double N = 34.986836655249725 / (pq * sqrt(den));   // 2 pi^(5/2)
            double Fm[L + 1];
            boys_fm<L>(alpha * (PQx * PQx + PQy * PQy + PQz * PQz), A.boys, Fm);
            const double m2a = -2.0 * alpha;
            double pw = N;
#pragma unroll
            for (int n = 0; n <= L; ++n) {
              s_F[lane][n] = Fm[n] * pw;
              if (n == L) {
                if (L == 0)
                  s_Rd[lane][0] = Fm[n] * pw;
                else
                  s_Rp[L % 2][lane][0] = Fm[n] * pw;
              }
              pw *= m2a;
            }
          }
          s_PQ[lane][0] = PQx;
          s_PQ[lane][1] = PQy;
          s_PQ[lane][2] = PQz;
          s_PQ[lane][3] = act ? 1.0 : 0.0;
        }
        // recursion R^(n)_{tuv} = coef R^(n+1)_{i2} + PQ_dir R^(n+1)_{i1}, levels L-1 .. 0
#pragma unroll
        for (int n = L - 1; n >= 0; --n) {
          __syncthreads();
          const int cnt = nherm(L - n);
          const int tot = nq * cnt;
          for (int e = lane; e < tot; e += 64) {
            const int ql = e / cnt, idx = e - ql * cnt;
            if (s_PQ[ql][3] == 0.0) continue;
            double r;
            if (idx == 0) {
              r = s_F[ql][n];
            } else {
              const int rec = s_rrec[idx];
              const int i1 = rec & 0x7ff, i2 = (rec >> 11) & 0x7ff, dir = (rec >> 22) & 3, coef = rec >> 24;
              const double *prv = s_Rp[(n + 1) % 2][ql];
              r = (double)coef * prv[i2] + s_PQ[ql][dir] * prv[i1];
            }
            if (n == 0)
              s_Rd[ql][s_dix[idx]] = r;
            else
              s_Rp[n % 2][ql][idx] = r;
          }
        }
        // stage-1 B operand: (-1)^{|k|} ET_ket(k, c) of the slot's ket primitive
        const bool kval = cval && kp < CD.nprim;
        const int kpc = kp < CD.nprim ? kp : CD.nprim - 1;
        const PrimRec PKm = kpc == 0 ? first_prim(CD) : A.prim[CD.prim0 + kpc];
        const double *etk = A.et + PKm.et;
        // loads from clamped addresses, masked afterwards (no branches)
        double eb[KSF];
#pragma unroll
        for (int kg = 0; kg < KSF; ++kg) {
          const int k = 4 * kg + lq;
          const bool ok = kval && k < NHF;
          const double v = etk[ok ? (long long)k * ldk + fc : 0];
          eb[kg] = ok ? (((ksgn >> kg) & 1u) ? -v : v) : 0.0;
        }
        __syncthreads();   // R tables complete
        // k step kg: NHG independent accumulators (a scheduling barrier per
        // step keeps the compiler from hoisting every read of the unrolled loop)
        // The step's offset passes through an empty asm statement so that its
        // NHG read addresses are formed in the step (one add each) instead of
        // all NHG x KSF of them being hoisted into registers.
        const double *Rq = s_Rd[myrt];
#pragma unroll
        for (int kg = 0; kg < KSF; ++kg) {
          int ko = koff[kg];
          asm volatile("" : "+v"(ko));
          const double *Rk = Rq + ko;
#pragma unroll
          for (int hg = 0; hg < NHG; ++hg) X[hg] = __builtin_amdgcn_mfma_f64_4x4x4f64(Rk[hoff[hg]], eb[kg], X[hg], 0, 0, 0);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      // stage 2, one row group of four bra rows at a time (the A operand's
      // row 4 rg + li): V(r, c) = sum_h ET_bra(h, r) X(h, c), two independent
      // accumulator chains (even and odd h steps); lane (q, b, i) then holds
      // V(4 rg + q, c0_b + i).  A contracted bra adds its later primitive
      // pairs to the stored block.
      for (int rg = 0; rg < nrg; ++rg) {
        const int r = 4 * rg + li;
        const int fr = pair_full(RB, A.sel, r < nr ? r : nr - 1);
        double v0 = 0.0, v1 = 0.0;
#pragma unroll
        for (int hg = 0; hg < NHG; ++hg) {
          const int h = 4 * hg + lq;
          const double e = etb[(h < NHE ? h : NHE - 1) * ldb + fr];
          if (hg & 1)
            v1 = __builtin_amdgcn_mfma_f64_4x4x4f64(h < NHE ? e : 0.0, X[hg], v1, 0, 0, 0);
          else
            v0 = __builtin_amdgcn_mfma_f64_4x4x4f64(h < NHE ? e : 0.0, X[hg], v0, 0, 0, 0);
        }
        const int ro = 4 * rg + lq;
        if (cval && ro < nr) {
          double *d = out + rs * ro + cs * c;
          *d = bp > 0 ? *d + (v0 + v1) : v0 + v1;
        }
      }
    }
  }
}

}  // namespace gi
