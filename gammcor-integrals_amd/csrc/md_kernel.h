// Hermite-matrix McMurchie-Davidson ERI engine for CDNA4 (gfx950).
//
// A shell quartet block is the matrix product
//   V(ab, cd) = sum_{h, k} E^AB(ab, h) (-1)^{|k|} R_{h+k}(alpha, P-Q) E^CD(cd, k)
// with h, k Hermite triples (t,u,v), E^AB the two-centre Hermite expansion
// coefficients of the bra pair (contraction coefficients, Gaussian product
// factor, solid-harmonic transform and AO normalization folded in) and
// R_{tuv} the Hermite Coulomb integrals (Helgaker, Joergensen, Olsen,
// "Molecular Electronic-Structure Theory", eqs. 9.5.33, 9.9.18-9.9.20).  The
// reference evaluates the same integrals through (e0|f0) blocks and the
// Head-Gordon-Pople horizontal recurrence (auto2e_WMatrix_part1.f90:99-148,
// Auto2e.py:298-338); the Boys function is evaluated the same way, once per
// primitive quartet at top order L = la+lb+lc+ld, then recursed downwards
// (boys.f90:202-281).
//
// Per shell pair the table ET(h, r) = E^AB(r, h) is built once per launch
// (k_build_et), r = full orbital-pair index a + na*b.  Per quartet one
// wavefront builds R_{h+k} in LDS and evaluates
//   stage 1  X(h, c) = sum_k (-1)^{|k|} R_{h+k} ET_ket(k, c)      (MFMA f64 16x16x4)
//   stage 2  V(r, c) = sum_h ET_bra(h, r) X(h, c)                (MFMA f64 16x16x4)
// sixteen ket columns c at a time.  Rows and columns are the destination's
// own orbital pairs (full block, same-shell lower triangle, or a selection
// list), so the result tiles go straight to the chol2_write_V* layouts of
// TwoStepCholesky_Step{1,2}.f90 with row-contiguous (coalesced) stores.
#pragma once
#include "gi_common.h"

namespace gi {

typedef double d4 __attribute__((ext_vector_type(4)));

// One shell pair as the caller sees it.  The destination address of
// element (jr, jc) of a quartet block is  out + row.off + col.off + jr + ldo * jc.
struct PairDesc {
  int s0, s1;       // shells (0-based), caller order
  int flags;        // PAIR_SAME: fold to lower triangle; PAIR_SEL: selection list
  int sel0, nsel;   // selection list slice: 0-based full-block indices a + na*b
  int pt;           // pair-table entry (filled by the launch layer)
  long long off;    // destination offset contributed by this pair
};
enum { PAIR_SAME = 1, PAIR_SEL = 2 };
enum { LAUNCH_PRODUCT = 0, LAUNCH_ZIP = 1, LAUNCH_DIAG = 2 };

// pair table
struct PairHdr {
  int n0, n1;       // functions per shell (spherical or Cartesian)
  int nprim;        // primitive pairs
  int prim0;        // first PrimRec
};
struct PrimRec {
  double px, py, pz, p;
  long long et;     // offset (doubles) of the ET block [nherm(l0+l1)][n0*n1]
  long long pad;
};
struct EtJob {
  int s0, s1, ia, ib;   // shells, primitive indices
  int prim, pad;        // PrimRec slot
  long long et;         // ET block offset
};
// Everything a kernel needs about one bra or ket pair of a launch, packed so
// that a quartet costs one independent load per side (k_pack_recs): the first
// primitive pair inline, the rest (contracted pairs) through prim0.
struct QRec {
  double px, py, pz, p;   // first primitive pair
  long long et;           // its ET block
  long long off;          // destination offset (PairDesc::off)
  int nrow, flags, sel0, n0;
  int n1, nprim, prim0, pad;
};
__host__ __device__ inline PrimRec first_prim(const QRec &q) { return PrimRec{q.px, q.py, q.pz, q.p, q.et, 0}; }

constexpr int LRMAX = 4 * MAXL;                 // highest R-table order
constexpr int NHRMAX = (LRMAX + 1) * (LRMAX + 2) * (LRMAX + 3) / 6;

// c_herm[idx]: packed (t,u,v) of Hermite index idx.
// c_rrec[idx]: recurrence of R_{tuv}, idx >= 1: i1 | i2 << 11 | dir << 22 | coef << 24,
//   R^(n)_{tuv} = coef R^(n+1)_{i2} + PQ_dir R^(n+1)_{i1}
__constant__ extern int c_herm[NHRMAX];
__constant__ extern int c_rrec[NHRMAX];
__constant__ extern double c_spher[SPHER_TOTAL];

__host__ __device__ constexpr int nsph(int l, int sph) { return (sph && l >= 2) ? 2 * l + 1 : ncart(l); }
__host__ __device__ constexpr int imax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int imin(int a, int b) { return a < b ? a : b; }

// ------------------------------------------------------------- Boys function
// F_0..F_M at x: 5-point Chebyshev interpolation of F_M on the interval of x,
// then downward recursion; x >= 25: 1/2 sqrt(pi/x) and upward recursion
// (boys.f90:202-281, chebinterp.f90:122-149).
template <int M>
__device__ __forceinline__ void boys_fm(double x, const double *__restrict__ tab, double (&F)[M + 1]) {
  constexpr double dx = BOYS_XASYM / (double)BOYS_NINT;
  constexpr double inv_dx = (double)BOYS_NINT / BOYS_XASYM;
  if (x < BOYS_XASYM) {
    int k = (int)floor(x / dx);   // same interval choice as floor(x / DELTAX)
    if (k > BOYS_NINT - 1) k = BOYS_NINT - 1;
    if (k < 0) k = 0;
    const double *c = tab + 5 * (k + BOYS_NINT * M);
    double ab_sum = (double)(2 * k + 1) * dx;
    double y = (2.0 * x - ab_sum) * inv_dx;
    double twoy = 2.0 * y;
    double d5 = c[4];
    double d4v = twoy * d5 + c[3];
    double d3 = twoy * d4v - d5 + c[2];
    double d2 = twoy * d3 - d4v + c[1];
    double f = y * d2 - d3 + 0.5 * c[0];
    F[M] = f;
    double expmx = exp(-x), twox = 2.0 * x;
#pragma unroll
    for (int l = M; l >= 1; --l) {
      f = (twox * f + expmx) * (1.0 / (double)(2 * l - 1));
      F[l - 1] = f;
    }
  } else {
    double f = 0.5 * sqrt(3.141592653589793 / x);
    F[0] = f;
    double expmx = exp(-x), beta = 1.0 / (2.0 * x);
#pragma unroll
    for (int l = 1; l <= M; ++l) {
      f = beta * ((double)(2 * l - 1) * f - expmx);
      F[l] = f;
    }
  }
}

// ------------------------------------------------------------- kernel shape
__host__ __device__ constexpr int nab_max(int E) {   // largest Cartesian pair block, l <= MAXL
  int m = 1;
  for (int l0 = 0; l0 <= MAXL; ++l0)
    for (int l1 = 0; l1 <= MAXL; ++l1)
      if (l0 + l1 == E) m = imax(m, ncart(l0) * ncart(l1));
  return m;
}

template <int E_, int F_>
struct MdShape {
  static constexpr int E = E_, F = F_, L = E + F;
  static constexpr int NHE = nherm(E), NHF = nherm(F), NHL = nherm(L);
  static constexpr int HT = (NHE + 15) / 16;    // stage-1 row tiles (h)
  static constexpr int KSF = (NHF + 3) / 4;     // stage-1 k steps
  static constexpr int KSE = (NHE + 3) / 4;     // stage-2 k steps
  static constexpr int RT = (nab_max(E) + 15) / 16;   // stage-2 column tiles (bra rows)
  static constexpr int NIDX = KSF * HT * 64;
};

struct MdArgs {
  const PrimRec *prim;
  const double *et;
  const QRec *rrec, *crec;       // packed bra / ket pairs of the launch
  long long nrows, ncols, ntiles;
  int mode, kt;                  // LAUNCH_*, ket pairs per tile (PRODUCT)
  int trans;                     // kernel bra side = caller's ket side (store transposed)
  int rowsum;                    // add each row's sum over the ket columns to out[off + r]
                                 // (atomically; PRODUCT, not transposed)
  double kappa;
  double *out;
  long long ldo;
  const int *sel;                // selection lists
  const double *boys;            // (5, BOYS_NINT, 0:BOYS_MAXM)
};

// full-block index a + n0 b of destination row r of a pair
__device__ __forceinline__ int pair_full(const QRec &d, const int *sel, int r) {
  if (d.flags & PAIR_SEL) return sel[d.sel0 + r];
  if (d.flags & PAIR_SAME) {
    // column-major lower triangle position r -> (a, b), a >= b
    const int n = d.n0;
    int b = 0, rem = r;
    while (rem >= n - b) { rem -= n - b; ++b; }
    return (b + rem) + n * b;
  }
  return r;
}

// R_{tuv} (times 2 pi^{5/2} / (p q sqrt(p + q + kappa p q))) for one primitive
// quartet into Ra (level 0); Rb is the ping-pong partner.  Both hold a zero in
// slot NHL.  Ends with a barrier.
template <int L>
__device__ __forceinline__ void rtable(double *Ra, double *Rb, const PrimRec &pb, const PrimRec &pk,
                                       double kappa, const double *__restrict__ boys, int lane,
                                       const int *__restrict__ rrec) {
  const double PQx = pb.px - pk.px, PQy = pb.py - pk.py, PQz = pb.pz - pk.pz;
  const double pq = pb.p * pk.p;
  const double den = pb.p + pk.p + kappa * pq;
  const double alpha = pq / den;
  const double N = 34.986836655249725 / (pq * sqrt(den));   // 2 pi^(5/2)
  double Fm[L + 1];
  boys_fm<L>(alpha * (PQx * PQx + PQy * PQy + PQz * PQz), boys, Fm);
  const double m2a = -2.0 * alpha;
  double pw = N;
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
#pragma unroll
    for (int base = 0; base < nherm(L); base += 64) {
      const int idx = base + lane;
      if (base < cnt && idx < cnt) {
        double r;
        if (idx == 0) {
          r = Fm[n];
        } else {
          const int rec = rrec[idx];
          const int i1 = rec & 0x7ff, i2 = (rec >> 11) & 0x7ff, dir = (rec >> 22) & 3, coef = rec >> 24;
          const double pqd = dir == 0 ? PQx : (dir == 1 ? PQy : PQz);
          r = (double)coef * prv[i2] + pqd * prv[i1];
        }
        cur[idx] = r;
      }
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------- quartet kernel
// One wavefront per tile of one bra pair and A.kt ket pairs.  Stage-1 results
// X(h, c) of successive ket columns (16 at a time) are appended to a ring of
// 32 columns in LDS; every 16 queued columns -- possibly from different ket
// pairs, which share ET_bra -- go through stage 2 together, so thin ket pairs
// (ss, ps, a few selected pivots) still fill the 16-wide MFMA tiles.  A
// contracted bra pair runs the whole tile once per bra primitive pair and
// accumulates into the destination.  In a transposed launch (A.trans) the
// kernel's bra side is the caller's ket side and rows/columns of the stored
// block are exchanged.
constexpr int MD_RING = 32;
constexpr int MD_NCV = 4;   // slices of at most this many columns take stage 1 on the VALU

template <int E, int F>
__global__ void __launch_bounds__(64) md_kernel(MdArgs A) {
  using S = MdShape<E, F>;
  constexpr int HT = S::HT, KSF = S::KSF, KSE = S::KSE, RT = S::RT, NHE = S::NHE, NHF = S::NHF;
  constexpr int NHL = S::NHL;
  // unroll factors of the k-step loops: full for small classes, partial where
  // hoisting every operand load of a fully unrolled loop would spill
  constexpr int U1 = KSF * HT > 24 ? 1 : KSF;
  constexpr int U2 = KSE * RT > 24 ? 1 : KSE;
  __shared__ unsigned short s_idx[S::NIDX];
  __shared__ double s_R[2][NHL + 1];
  __shared__ double s_X[HT * 16][MD_RING];
  __shared__ long long s_qbase[MD_RING];
  __shared__ double s_E[NHF][MD_NCV];
  __shared__ int s_rrec[NHL];   // R recurrence codes, read by every R table of the launch

  const int lane = threadIdx.x;
  const int i16 = lane & 15, g = lane >> 4;
  // stage-1 operand map: lane (i, kk) of step ks, row tile ht reads
  // (-1)^{|k|} R_{h+k}, h = 16 ht + i, k = 4 ks + kk; slot NHL (zero) pads.
  for (int e = lane; e < S::NIDX; e += 64) {
    const int ln = e & 63, blk = e >> 6;
    const int ks = blk / HT, ht = blk - ks * HT;
    const int h = 16 * ht + (ln & 15), k = 4 * ks + (ln >> 4);
    unsigned short v = (unsigned short)NHL;
    if (h < NHE && k < NHF) {
      const int th = c_herm[h], tk = c_herm[k];
      const int t = px(th) + px(tk), u = py(th) + py(tk), w = pz(th) + pz(tk);
      v = (unsigned short)hidx(t, u, w);
      if ((px(tk) + py(tk) + pz(tk)) & 1) v |= 0x8000;
    }
    s_idx[e] = v;
  }
  for (int e = lane; e < HT * 16 * MD_RING; e += 64) (&s_X[0][0])[e] = 0.0;
  for (int e = lane; e < NHL; e += 64) s_rrec[e] = c_rrec[e];
  if (lane == 0) {
    s_R[0][NHL] = 0.0;
    s_R[1][NHL] = 0.0;
  }
  __syncthreads();

  // XCD-aware order: consecutive tiles (same bra pair) on one XCD's L2
  const int G = gridDim.x;
  const int w = blockIdx.x;
  const long long v0 = (G % 8 == 0) ? (long long)(w % 8) * (G / 8) + w / 8 : w;
  const long long nct = (A.mode == LAUNCH_PRODUCT) ? (A.ncols + A.kt - 1) / A.kt : 1;
  const long long rs = A.trans ? A.ldo : 1;   // destination stride of a kernel row

  for (long long t = v0; t < A.ntiles; t += G) {
    long long ir, jc0;
    int nk;
    if (A.mode == LAUNCH_PRODUCT) {
      ir = t / nct;
      jc0 = (t - ir * nct) * A.kt;
      nk = (int)((A.ncols - jc0) < A.kt ? (A.ncols - jc0) : A.kt);
    } else {
      ir = t;
      jc0 = t;
      nk = 1;
    }
    const QRec RB = A.rrec[ir];
    const int nr = RB.nrow;
    const int ldb = RB.n0 * RB.n1;
    int fr[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
      const int r = 16 * rt + i16;
      fr[rt] = pair_full(RB, A.sel, r < nr ? r : nr - 1);
    }
    const int nrt = (nr + 15) / 16;
    double *const out = A.out + RB.off;

    for (int bp = 0; bp < RB.nprim; ++bp) {
      const PrimRec PB = bp == 0 ? first_prim(RB) : A.prim[RB.prim0 + bp];
      const double *etb = A.et + PB.et;
      const bool accumulate = bp > 0;
      int head = 0, q = 0;   // ring of queued stage-1 columns

      // stage 2 on the n oldest queued columns, stored (or added) to out
      auto flush = [&](int n) {
        d4 vacc[RT];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) vacc[rt] = d4{0.0, 0.0, 0.0, 0.0};
        const int col = (head + i16) & (MD_RING - 1);
#pragma unroll U2
        for (int ks = 0; ks < KSE; ++ks) {
          const int h = 4 * ks + g;
          const double a = s_X[h][col];
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) {
            if (rt < nrt) {
              const double b = (h < NHE) ? etb[(long long)h * ldb + fr[rt]] : 0.0;
              vacc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, vacc[rt], 0, 0, 0);
            }
          }
        }
        // lane (j, g) holds V(row 16 rt + j, queued column 4 qq + g)
        if (A.rowsum) {
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) {
            double sv = 0.0;
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) sv += (4 * qq + g < n) ? vacc[rt][qq] : 0.0;
            sv += __shfl_xor(sv, 16);
            sv += __shfl_xor(sv, 32);
            const int r = 16 * rt + i16;
            if (g == 0 && rt < nrt && r < nr) atomicAdd(out + r, sv);
          }
          head = (head + n) & (MD_RING - 1);
          q -= n;
          __syncthreads();   // ring slots are free again
          return;
        }
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
          const int ci = 4 * qq + g;
          if (ci < n) {
            const long long base = s_qbase[(head + ci) & (MD_RING - 1)];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
              const int r = 16 * rt + i16;
              if (rt < nrt && r < nr) {
                double *d = out + base + rs * r;
                *d = accumulate ? *d + vacc[rt][qq] : vacc[rt][qq];
              }
            }
          }
        }
        head = (head + n) & (MD_RING - 1);
        q -= n;
        __syncthreads();   // ring slots are free again
      };

      for (int kq = 0; kq < nk; ++kq) {
        const QRec CD = A.crec[jc0 + kq];
        const int nc = CD.nrow;
        const int ldk = CD.n0 * CD.n1;
        const bool single = (CD.nprim == 1);
        if (single) rtable<E + F>(s_R[0], s_R[1], PB, first_prim(CD), A.kappa, A.boys, lane, s_rrec);

        for (int c0 = 0; c0 < nc; c0 += 16) {
          const int cc = c0 + i16;
          const int ncs = (nc - c0) < 16 ? (nc - c0) : 16;
          if (ncs <= MD_NCV && A.mode != LAUNCH_DIAG) {
            // narrow slice: X(h, c) = sum_k (-1)^{|k|} R_{h+k} ET_ket(k, c) on the
            // VALU, lane h (and h + 64), the slice of ET_ket staged in LDS
            double xv[2][MD_NCV];
#pragma unroll
            for (int ps = 0; ps < 2; ++ps)
#pragma unroll
              for (int c = 0; c < MD_NCV; ++c) xv[ps][c] = 0.0;
            for (int kp = 0; kp < CD.nprim; ++kp) {
              const PrimRec PK = kp == 0 ? first_prim(CD) : A.prim[CD.prim0 + kp];
              if (!single) rtable<E + F>(s_R[0], s_R[1], PB, PK, A.kappa, A.boys, lane, s_rrec);
              const double *etk = A.et + PK.et;
              for (int e = lane; e < NHF * MD_NCV; e += 64) {
                const int k = e / MD_NCV, c = e - k * MD_NCV;
                s_E[k][c] = (c < ncs) ? etk[(long long)k * ldk + pair_full(CD, A.sel, c0 + c)] : 0.0;
              }
              __syncthreads();
              const double *Rf = s_R[0];
#pragma unroll
              for (int ps = 0; ps < 2; ++ps) {
                const int h = lane + 64 * ps;
                if (h < NHE) {
#pragma unroll 4
                  for (int k = 0; k < NHF; ++k) {
                    const unsigned short e = s_idx[((k >> 2) * HT + (h >> 4)) * 64 + (h & 15) + 16 * (k & 3)];
                    double r = Rf[e & 0x7fff];
                    r = (e & 0x8000) ? -r : r;
#pragma unroll
                    for (int c = 0; c < MD_NCV; ++c) xv[ps][c] += r * s_E[k][c];
                  }
                }
              }
              __syncthreads();
            }
#pragma unroll
            for (int c = 0; c < MD_NCV; ++c) {
              if (c < ncs) {
                const int slot = (head + q + c) & (MD_RING - 1);
#pragma unroll
                for (int ps = 0; ps < 2; ++ps) {
                  const int h = lane + 64 * ps;
                  if (h < HT * 16) s_X[h][slot] = h < NHE ? xv[ps][c] : 0.0;
                }
                if (lane == 0) s_qbase[slot] = A.trans ? CD.off + (c0 + c) : CD.off + A.ldo * (c0 + c);
              }
            }
            q += ncs;
            __syncthreads();
            if (q >= 16) flush(16);
            continue;
          }
          const int fc = pair_full(CD, A.sel, cc < nc ? cc : nc - 1);
          d4 xacc[HT];
#pragma unroll
          for (int ht = 0; ht < HT; ++ht) xacc[ht] = d4{0.0, 0.0, 0.0, 0.0};
          for (int kp = 0; kp < CD.nprim; ++kp) {
            const PrimRec PK = kp == 0 ? first_prim(CD) : A.prim[CD.prim0 + kp];
            if (!single) rtable<E + F>(s_R[0], s_R[1], PB, PK, A.kappa, A.boys, lane, s_rrec);
            const double *etk = A.et + PK.et;
            const double *Rf = s_R[0];
#pragma unroll U1
            for (int ks = 0; ks < KSF; ++ks) {
              const int k = 4 * ks + g;
              const double b = (k < NHF) ? etk[(long long)k * ldk + fc] : 0.0;
#pragma unroll
              for (int ht = 0; ht < HT; ++ht) {
                const unsigned short e = s_idx[(ks * HT + ht) * 64 + lane];
                double a = Rf[e & 0x7fff];
                a = (e & 0x8000) ? -a : a;
                xacc[ht] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, xacc[ht], 0, 0, 0);
              }
            }
          }
          if (A.mode == LAUNCH_DIAG) {
            // diagonal elements only: V(c, c) = sum_h ET_bra(h, c) X(h, c)
            __syncthreads();
#pragma unroll
            for (int ht = 0; ht < HT; ++ht)
#pragma unroll
              for (int qq = 0; qq < 4; ++qq) s_X[16 * ht + 4 * qq + g][i16] = xacc[ht][qq];
            __syncthreads();
            if (g == 0 && cc < nc) {
              double s = 0.0;
#pragma unroll 4
              for (int h = 0; h < NHE; ++h) s += etb[(long long)h * ldb + fc] * s_X[h][i16];
              double *d = out + CD.off + cc;
              *d = accumulate ? *d + s : s;
            }
            __syncthreads();
            continue;
          }
          // append the slice's columns to the ring.  MFMA f64 16x16x4 result
          // layout: element (i, j) of a tile sits in lane j + 16 (i % 4),
          // register i / 4, so lane (j, g) holds rows 16 ht + 4 qq + g of column j.
          const int slot = (head + q + i16) & (MD_RING - 1);
          if (i16 < ncs) {
#pragma unroll
            for (int ht = 0; ht < HT; ++ht)
#pragma unroll
              for (int qq = 0; qq < 4; ++qq) s_X[16 * ht + 4 * qq + g][slot] = xacc[ht][qq];
            if (g == 0) s_qbase[slot] = A.trans ? CD.off + cc : CD.off + A.ldo * cc;
          }
          q += ncs;
          __syncthreads();
          if (q >= 16) flush(16);
        }
      }
      if (q > 0) flush(q);
    }
  }
}

// ------------------------------------------------------------- lane kernel
// Small and thin-bra classes (the list in md_classes.inc; uncontracted pairs): one lane per
// quartet, everything in registers.  A wave takes 64 consecutive bra pairs of
// the row list against one ket pair (PRODUCT; the ket side is then uniform
// across the wave and its loads scalar), or 64 consecutive zipped pairs.
//   y(k)    = sum_h ET_bra(h, r) R_{h+k}
//   V(r, c) = sum_k (-1)^{|k|} y(k) ET_ket(k, c)
constexpr int LANE_EMAX = 4;   // largest E or F with a lane kernel (md_classes.inc lists them)

__host__ __device__ constexpr int hsum(int h, int k) {
  return hidx(px(herm_of(h)) + px(herm_of(k)), py(herm_of(h)) + py(herm_of(k)), pz(herm_of(h)) + pz(herm_of(k)));
}
__host__ __device__ constexpr int hord(int h) { return px(herm_of(h)) + py(herm_of(h)) + pz(herm_of(h)); }

// compile-time index tables (so that unrolled register code has constant indices)
template <int E, int F>
struct SumTab {
  int v[nherm(E)][nherm(F)];
  int odd[nherm(F)];
  constexpr SumTab() : v(), odd() {
    for (int h = 0; h < nherm(E); ++h)
      for (int k = 0; k < nherm(F); ++k) v[h][k] = hsum(h, k);
    for (int k = 0; k < nherm(F); ++k) odd[k] = hord(k) & 1;
  }
};
template <int L>
struct RecTab {
  int i1[nherm(L)], i2[nherm(L)], dir[nherm(L)], c[nherm(L)];
  constexpr RecTab() : i1(), i2(), dir(), c() {
    for (int idx = 1; idx < nherm(L); ++idx) {
      const int p = herm_of(idx);
      const int t = px(p), u = py(p), v = pz(p);
      const int d = t > 0 ? 0 : (u > 0 ? 1 : 2);
      dir[idx] = d;
      c[idx] = d == 0 ? t : (d == 1 ? u : v);
      i1[idx] = hidx(t - (d == 0), u - (d == 1), v - (d == 2));
      i2[idx] = c[idx] >= 2 ? hidx(t - 2 * (d == 0), u - 2 * (d == 1), v - 2 * (d == 2)) : 0;
    }
  }
};

// R_{tuv} of one primitive quartet in registers, recursion in place: level n
// entries in decreasing index order only read lower indices of level n + 1.
template <int L>
__device__ __forceinline__ void rtable_reg(double (&R)[nherm(L)], const PrimRec &pb, const PrimRec &pk,
                                           double kappa, const double *__restrict__ boys) {
  const double PQ[3] = {pb.px - pk.px, pb.py - pk.py, pb.pz - pk.pz};
  const double pq = pb.p * pk.p;
  const double den = pb.p + pk.p + kappa * pq;
  const double alpha = pq / den;
  const double N = 34.986836655249725 / (pq * sqrt(den));   // 2 pi^(5/2)
  double Fm[L + 1];
  boys_fm<L>(alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]), boys, Fm);
  const double m2a = -2.0 * alpha;
  double pw = N;
#pragma unroll
  for (int n = 0; n <= L; ++n) {
    Fm[n] *= pw;
    pw *= m2a;
  }
  constexpr RecTab<L> T{};
  R[0] = Fm[L];
#pragma unroll
  for (int n = L - 1; n >= 0; --n) {
#pragma unroll
    for (int idx = nherm(L - n) - 1; idx >= 1; --idx) {
      double r = PQ[T.dir[idx]] * R[T.i1[idx]];
      if (T.c[idx] >= 2) r += (double)(T.c[idx] - 1) * R[T.i2[idx]];
      R[idx] = r;
    }
    R[0] = Fm[n];
  }
}

template <int E, int F>
__global__ void __launch_bounds__(256) md_lane_kernel(MdArgs A) {
  constexpr int NHE = nherm(E), NHF = nherm(F), L = E + F;
  constexpr int NRM = nab_max(E), NCM = nab_max(F);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long nrg = (A.mode == LAUNCH_PRODUCT) ? (A.nrows + 63) / 64 : (A.nrows + 63) / 64;
  const long long ntile = (A.mode == LAUNCH_PRODUCT) ? nrg * A.ncols : nrg;
  for (long long t = (long long)blockIdx.x * 4 + wave; t < ntile; t += (long long)gridDim.x * 4) {
    long long ir, jc;
    if (A.mode == LAUNCH_PRODUCT) {
      const long long rg = t / A.ncols;
      jc = t - rg * A.ncols;
      ir = rg * 64 + lane;
    } else {
      ir = t * 64 + lane;
      jc = ir;
    }
    if (ir >= A.nrows) continue;
    const QRec RD = A.rrec[ir];
    const QRec CD = A.crec[jc];
    const PrimRec PB = first_prim(RD), PK = first_prim(CD);
    const int nr = RD.nrow, nc = CD.nrow;
    const int ldb = RD.n0 * RD.n1, ldk = CD.n0 * CD.n1;
    const double *etb = A.et + PB.et, *etk = A.et + PK.et;
    double *dst = A.out + RD.off + CD.off;
    // transposed launch: the kernel's bra rows are the destination's columns
    const long long rs = A.trans ? A.ldo : 1, cs = A.trans ? 1 : A.ldo;
    double R[nherm(L)];
    rtable_reg<L>(R, PB, PK, A.kappa, A.boys);
#pragma unroll 1
    for (int r = 0; r < NRM; ++r) {
      if (r < nr) {
        const int fr = pair_full(RD, A.sel, r);
        double eb[NHE];
#pragma unroll
        for (int h = 0; h < NHE; ++h) eb[h] = etb[(long long)h * ldb + fr];
        constexpr SumTab<E, F> ST{};
        double y[NHF];
#pragma unroll
        for (int k = 0; k < NHF; ++k) {
          double s = 0.0;
#pragma unroll
          for (int h = 0; h < NHE; ++h) s += eb[h] * R[ST.v[h][k]];
          y[k] = ST.odd[k] ? -s : s;
        }
        if (A.mode == LAUNCH_DIAG) {
          // diagonal element (r, r) only
          double v = 0.0;
#pragma unroll
          for (int k = 0; k < NHF; ++k) v += y[k] * etk[(long long)k * ldk + fr];
          dst[r] = v;
        } else {
#pragma unroll
          for (int c = 0; c < NCM; ++c) {
            if (c < nc) {
              const int fc = pair_full(CD, A.sel, c);
              double v = 0.0;
#pragma unroll
              for (int k = 0; k < NHF; ++k) v += y[k] * etk[(long long)k * ldk + fc];
              dst[rs * r + cs * c] = v;
            }
          }
        }
      }
    }
  }
}

}  // namespace gi
