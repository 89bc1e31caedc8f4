// ERI kernels for one canonical angular-momentum class (LA>=LB, LC>=LD,
// (LA,LB)>=(LC,LD)), McMurchie-Davidson (e0|f0) blocks + Head-Gordon-Pople
// horizontal recurrence, with the Boys function evaluated once per (e,f)
// block at top order e+f and recursed downwards, as the reference's generated
// engine does (auto2e_WMatrix_part1.f90:99-148,445-506; Auto2e.py:298-338).
//
// A "group" of G lanes evaluates one shell quartet; a 64-lane workgroup holds
// 64/G groups.  Per-quartet working arrays live in registers (G == 1, fully
// unrolled), in LDS, or in a global workspace for the largest classes.
//
// Work list: a launch covers one canonical class and is either the Cartesian
// product of a list of bra shell pairs ("rows") and a list of ket shell pairs
// ("columns"), or the element-wise pairing of two such lists.  All bra pairs
// of a launch share the caller-order angular momenta (la, lb), all ket pairs
// (lc, ld), so the permutation from the caller's (A,B|C,D) to canonical order
// is a launch constant.  Results are scattered straight into the destination
// matrix (rows = bra orbital pairs, columns = ket orbital pairs, same-shell
// pairs folded to a lower triangle, optional selection lists), i.e. the
// chol2_write_V* layouts of TwoStepCholesky_Step{1,2}.f90.
#pragma once
#include "gi_common.h"

namespace gi {

// One shell pair as the caller sees it.  The destination address of
// element (jr, jc) of a quartet block is
//   out + row.off + col.off + jr + ldo * jc
struct PairDesc {
  int s0, s1;       // shells (0-based), caller order
  int flags;        // PAIR_SAME: fold to lower triangle; PAIR_SEL: selection list
  int sel0, nsel;   // selection list slice: 0-based full-block indices a + na*b
  int pad;
  long long off;    // destination offset contributed by this pair
};
enum { PAIR_SAME = 1, PAIR_SEL = 2 };
enum { LAUNCH_PRODUCT = 0, LAUNCH_ZIP = 1, LAUNCH_DIAG = 2 };

struct EriArgs {
  const DevShell *shells;
  const double *exps, *cntr, *norms;
  const double *boys;            // (5, BOYS_NINT, 0:BOYS_MAXM)
  const PairDesc *rdesc, *cdesc;
  const int *rows, *cols;        // indices into rdesc / cdesc
  long long nrows, ncols;
  int mode;                      // LAUNCH_*
  int perm;                      // caller slot of canonical position k in bits 2k..2k+1
  int spher;
  double kappa;
  double *out;
  long long ldo;
  const int *sel;                // selection lists
  double *gws;                   // global workspace (GLB mode)
  int npa, npb, npc, npd;        // primitive counts, uniform over the launch
};

__constant__ extern double c_spher[SPHER_TOTAL];

// ------------------------------------------------------------- class plan
__host__ __device__ constexpr int sum_ncart(int l0, int l1) {
  int s = 0;
  for (int l = l0; l <= l1; ++l) s += ncart(l);
  return s;
}
__host__ __device__ constexpr int tri(int n) { return n * (n + 1) / 2; }
__host__ __device__ constexpr int imax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int imin(int a, int b) { return a < b ? a : b; }
// largest HRR level buffer (per row) for transferring l2 onto l1
__host__ __device__ constexpr int hrr_lev(int l1, int l2) {
  int m = sum_ncart(l1, l1 + l2);
  for (int dl = 1; dl <= l2; ++dl) m = imax(m, sum_ncart(l1, l1 + l2 - dl) * ncart(dl));
  return m;
}
__host__ __device__ constexpr int nsph(int l, int sph) { return (sph && l >= 2) ? 2 * l + 1 : ncart(l); }

enum { MODE_REG = 0, MODE_LDS = 1, MODE_GLB = 2 };

template <int LA_, int LB_, int LC_, int LD_>
struct EriPlan {
  static constexpr int LA = LA_, LB = LB_, LC = LC_, LD = LD_;
  static constexpr int L = LA + LB + LC + LD;
  static constexpr int EMAX = LA + LB, FMAX = LC + LD;
  static constexpr int NR = sum_ncart(LA, LA + LB);
  static constexpr int NC = sum_ncart(LC, LC + LD);
  static constexpr int NA = ncart(LA), NB = ncart(LB), NCC = ncart(LC), ND = ncart(LD);
  static constexpr int NAB = NA * NB, NCD = NCC * ND;   // Cartesian (upper bounds)
  // phase 1 scratch: R ping-pong, S (all ket comps of the largest block), 1D E tables
  static constexpr int P1 = 2 * nherm(L) + ncart(FMAX) * nherm(EMAX) + 3 * tri(EMAX + 1) + 3 * tri(FMAX + 1);
  // ket HRR chunk of RC rows, bra HRR chunk of KC columns
  static constexpr int KLEV = imax(hrr_lev(LC, LD), NCD);
  static constexpr int BLEV = imax(hrr_lev(LA, LB), NAB);
  static constexpr int WSZ = imax(NR * NC, NAB * NCD);   // W, later overlaid by VS
  static constexpr int TSZ = NR * NCD;                    // TS
  static constexpr int BASE_NOCHUNK = WSZ + TSZ;
  // register mode if everything is tiny
  static constexpr int P23_FULL = 2 * imax(NR * KLEV, NCD * BLEV) + imax(NR * NCD, NCD * NAB);
  static constexpr int PER_Q_REG = BASE_NOCHUNK + imax(P1, P23_FULL);
  static constexpr bool REG = PER_Q_REG <= 112 && L <= 4;
  static constexpr int RC = REG ? NR : imin(NR, 8);
  static constexpr int KC = REG ? NCD : imin(NCD, 8);
  static constexpr int P23 = 2 * imax(RC * KLEV, KC * BLEV) + imax(RC * NCD, KC * NAB);
  static constexpr int PER_Q = BASE_NOCHUNK + imax(P1, P23);
  static constexpr int LDS_BUDGET = 32768 / 8;   // doubles per workgroup
  static constexpr int G_LDS = (PER_Q * 16 <= LDS_BUDGET) ? 4
                             : (PER_Q * 8 <= LDS_BUDGET) ? 8
                             : (PER_Q * 4 <= LDS_BUDGET) ? 16
                             : (PER_Q * 2 <= LDS_BUDGET) ? 32
                             : (PER_Q <= 2 * LDS_BUDGET) ? 64 : 0;
  static constexpr int MODE = REG ? MODE_REG : (G_LDS > 0 ? MODE_LDS : MODE_GLB);
  static constexpr int G = REG ? 1 : (G_LDS > 0 ? G_LDS : 64);
  static constexpr int QPW = 64 / G;
  static constexpr int LDS_DOUBLES = (MODE == MODE_LDS) ? QPW * PER_Q : 1;
};

// ------------------------------------------------------------- device pieces
// Boys F_0..F_M at x: 5-point Chebyshev interpolation of F_M on the interval
// of x, then downward recursion; x >= 25: 1/2 sqrt(pi/x) and upward
// recursion (boys.f90:202-281, chebinterp.f90:122-149).
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
    double d4 = twoy * d5 + c[3];
    double d3 = twoy * d4 - d5 + c[2];
    double d2 = twoy * d3 - d4 + c[1];
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

// 1D one-centre Hermite coefficients, packed E[i(i+1)/2 + t], i <= IM:
// E^i_0 = X E^{i-1}_0 + E^{i-1}_1, E^i_t = i/(2pt) E^{i-1}_{t-1}
// (auto2e_Hermite.f90:10-60)
template <int IM>
__device__ __forceinline__ void etab(double *E, double inv2p, double X, double e000) {
  E[0] = e000;
#pragma unroll
  for (int i = 1; i <= IM; ++i) {
    E[tri(i)] = X * E[tri(i - 1)] + (i >= 2 ? E[tri(i - 1) + 1] : 0.0);
#pragma unroll
    for (int t = 1; t <= i; ++t) E[tri(i) + t] = ((double)i / (double)t) * inv2p * E[tri(i - 1) + t - 1];
  }
}

template <int G>
__device__ __forceinline__ void gsync() {
  if constexpr (G > 1) __syncthreads();
}

// One (e,f) block for one primitive quartet: Boys, R_tuv, ket transform,
// bra transform, accumulated into W (row-major, row stride NC).
template <class P, int E, int F, int G>
__device__ __forceinline__ void mdblock(double *W, double *scr, int g, double alpha,
                                        const double (&Rpq)[3], const double (&Rpa)[3],
                                        const double (&Rqc)[3], double inv2p, double inv2q,
                                        double e000, const double *__restrict__ boys) {
  constexpr int LEF = E + F;
  constexpr int NH = nherm(LEF);
  double *R0 = scr, *R1 = scr + nherm(P::L);
  double *S = scr + 2 * nherm(P::L);
  double *Eb = S + ncart(P::FMAX) * nherm(P::EMAX);
  double *Ek = Eb + 3 * tri(P::EMAX + 1);
  constexpr int TE = tri(E + 1), TF = tri(F + 1);
  // 1D coefficient tables: bra x,y,z then ket x,y,z
  for (int d = g; d < 6; d += G) {
    if (d == 0) etab<E>(Eb, inv2p, Rpa[0], e000);
    else if (d == 1) etab<E>(Eb + TE, inv2p, Rpa[1], 1.0);
    else if (d == 2) etab<E>(Eb + 2 * TE, inv2p, Rpa[2], 1.0);
    else if (d == 3) etab<F>(Ek, inv2q, Rqc[0], 1.0);
    else if (d == 4) etab<F>(Ek + TF, inv2q, Rqc[1], 1.0);
    else etab<F>(Ek + 2 * TF, inv2q, Rqc[2], 1.0);
  }
  double Fm[LEF + 1];
  boys_fm<LEF>(alpha * (Rpq[0] * Rpq[0] + Rpq[1] * Rpq[1] + Rpq[2] * Rpq[2]), boys, Fm);
  const double m2a = -2.0 * alpha;
  // R^{(n)}_{tuv}, n = LEF..0 (Helgaker 9.9.18-20).  Ping-pong so that level 0
  // ends in R0.
  double *cur = (LEF % 2 == 0) ? R0 : R1;
  double *prv = (LEF % 2 == 0) ? R1 : R0;
  gsync<G>();
  if (g == 0) {
    double pw = 1.0;
#pragma unroll
    for (int i = 0; i < LEF; ++i) pw *= m2a;
    cur[0] = pw * Fm[LEF];
  }
#pragma unroll
  for (int n = LEF - 1; n >= 0; --n) {
    double *tmp = prv; prv = cur; cur = tmp;
    gsync<G>();
    double pw = 1.0;
#pragma unroll
    for (int i = 0; i < n; ++i) pw *= m2a;
    const int cnt = nherm(LEF - n);
#pragma unroll
    for (int idx = g; idx < cnt; idx += G) {
      const int p = herm_of(idx);
      const int t = px(p), u = py(p), v = pz(p);
      double r;
      if (t >= 1)
        r = (t >= 2 ? (double)(t - 1) * prv[hidx(t - 2, u, v)] : 0.0) + Rpq[0] * prv[hidx(t - 1, u, v)];
      else if (u >= 1)
        r = (u >= 2 ? (double)(u - 1) * prv[hidx(t, u - 2, v)] : 0.0) + Rpq[1] * prv[hidx(t, u - 1, v)];
      else if (v >= 1)
        r = (v >= 2 ? (double)(v - 1) * prv[hidx(t, u, v - 2)] : 0.0) + Rpq[2] * prv[hidx(t, u, v - 1)];
      else
        r = pw * Fm[n];
      cur[idx] = r;
    }
  }
  gsync<G>();
  // ket transform S[kc][h], h = hidx(t,u,v), t+u+v <= E
  constexpr int NHE = nherm(E);
  constexpr int NKC = ncart(F);
#pragma unroll
  for (int idx = g; idx < NKC * NHE; idx += G) {
    const int kc = idx / NHE, h = idx - kc * NHE;
    const int kp = cart_of(F, kc), hp = herm_of(h);
    const int kx = px(kp), ky = py(kp), kz = pz(kp);
    const int t = px(hp), u = py(hp), v = pz(hp);
    double s = 0.0;
    for (int ph = 0; ph <= kz; ++ph) {
      const double ez = Ek[2 * TF + tri(kz) + ph];
      for (int nu = 0; nu <= ky; ++nu) {
        const double eyz = Ek[TF + tri(ky) + nu] * ez;
        for (int ta = 0; ta <= kx; ++ta) {
          const double c = Ek[tri(kx) + ta] * eyz * R0[hidx(t + ta, u + nu, v + ph)];
          s = ((ta + nu + ph) & 1) ? s - c : s + c;
        }
      }
    }
    S[kc * NHE + h] = s;
  }
  gsync<G>();
  // bra transform into W rows of level E, columns of level F
  constexpr int ROFF = sum_ncart(P::LA, E - 1);
  constexpr int COFF = sum_ncart(P::LC, F - 1);
  constexpr int NEC = ncart(E);
#pragma unroll
  for (int idx = g; idx < NEC * NKC; idx += G) {
    const int kc = idx / NEC, ec = idx - kc * NEC;
    const int ep = cart_of(E, ec);
    const int ex = px(ep), ey = py(ep), ez = pz(ep);
    double w = 0.0;
    for (int v = 0; v <= ez; ++v) {
      const double bz = Eb[2 * TE + tri(ez) + v];
      for (int u = 0; u <= ey; ++u) {
        const double byz = Eb[TE + tri(ey) + u] * bz;
        for (int t = 0; t <= ex; ++t) w += Eb[tri(ex) + t] * byz * S[kc * NHE + hidx(t, u, v)];
      }
    }
    W[(ROFF + ec) * P::NC + COFF + kc] += w;
  }
  gsync<G>();
}

// compile-time loop over the (e,f) blocks
template <class P, int KB, int KD, int G>
__device__ __forceinline__ void all_blocks(double *W, double *scr, int g, double alpha,
                                           const double (&Rpq)[3], const double (&Rpa)[3],
                                           const double (&Rqc)[3], double inv2p, double inv2q,
                                           double e000, const double *__restrict__ boys) {
  if constexpr (KB <= P::LB) {
    if constexpr (KD <= P::LD) {
      mdblock<P, P::LA + KB, P::LC + KD, G>(W, scr, g, alpha, Rpq, Rpa, Rqc, inv2p, inv2q, e000, boys);
      all_blocks<P, KB, KD + 1, G>(W, scr, g, alpha, Rpq, Rpa, Rqc, inv2p, inv2q, e000, boys);
    } else {
      all_blocks<P, KB + 1, 0, G>(W, scr, g, alpha, Rpq, Rpa, Rqc, inv2p, inv2q, e000, boys);
    }
  }
}

// HRR on a chunk of vectors.  Input vector j, packed level index k (levels
// L1..L1+L2) sits at in[j*ist + k*iel]; output vector j, index c + ncart(L1)*d
// at out[j*ost + (c + ncart(L1)*d)].  tmp holds 2*nvec*hrr_lev(L1,L2) doubles.
// (c, d+1_i) = (c+1_i, d) + X_i (c, d), i = first nonzero of d (x,y,z).
template <int L1, int L2, int G>
__device__ __forceinline__ void hrr_chunk(const double *in, int ist, int iel, double *out, int ost,
                                          double *tmp, int nvec, const double (&X)[3], int g) {
  constexpr int LEV = hrr_lev(L1, L2);
  const double *prev = in;
  int pst = ist, pel = iel, pnpk = sum_ncart(L1, L1 + L2);
#pragma unroll
  for (int dl = 1; dl <= L2; ++dl) {
    constexpr int dummy = 0;
    (void)dummy;
    const int npk = sum_ncart(L1, L1 + L2 - dl);
    const int nd = ncart(dl);
    double *cur = (dl == L2) ? nullptr : tmp + ((dl & 1) ? 0 : nvec * LEV);
    const int cst = (dl == L2) ? ost : npk * nd;
    double *dst = (dl == L2) ? out : cur;
    const int tot = nvec * npk * nd;
    for (int idx = g; idx < tot; idx += G) {
      const int j = idx / (npk * nd);
      const int rem = idx - j * (npk * nd);
      const int d = rem / npk, cp = rem - d * npk;
      const int dp = cart_of(dl, d);
      const int dx = px(dp), dy = py(dp), dz = pz(dp);
      const int dir = dx > 0 ? 0 : (dy > 0 ? 1 : 2);
      const int dm = cidx(dx - (dir == 0), dy - (dir == 1), dz - (dir == 2));
      // packed c' -> (level, component)
      int lv = L1, base = 0;
      while (cp - base >= ncart(lv)) { base += ncart(lv); ++lv; }
      const int cc = cart_of(lv, cp - base);
      const int cpp = base + ncart(lv) + cidx(px(cc) + (dir == 0), py(cc) + (dir == 1), pz(cc) + (dir == 2));
      const double *pv = prev + j * pst;
      const double v1 = pv[(cpp + pnpk * dm) * pel];
      const double v0 = pv[(cp + pnpk * dm) * pel];
      dst[j * cst + cp + npk * d] = v1 + X[dir] * v0;
    }
    gsync<G>();
    prev = cur;
    pst = npk * nd;
    pel = 1;
    pnpk = npk;
  }
  if constexpr (L2 == 0) {
    constexpr int n1 = ncart(L1);
    for (int idx = g; idx < nvec * n1; idx += G) {
      const int j = idx / n1, c = idx - j * n1;
      out[j * ost + c] = in[j * ist + c * iel];
    }
    gsync<G>();
  }
}

// apply the solid-harmonic transform to the "slow" (d) and "fast" (c) index of
// vectors v[j*st + c + n1*d], c < ncart(L1), d < ncart(L2); result written to
// out[j*ost + cs + ns1*ds].  tmp: nvec*ncart(L1)*nsph(L2) doubles.
template <int L1, int L2, int G>
__device__ __forceinline__ void sph_chunk(const double *v, int st, double *out, int ost, double *tmp,
                                          int nvec, int sph, int g) {
  constexpr int n1 = ncart(L1), n2 = ncart(L2);
  const int s1 = sph && L1 >= 2 ? 2 * L1 + 1 : n1;
  const int s2 = sph && L2 >= 2 ? 2 * L2 + 1 : n2;
  const double *U1 = c_spher + spher_off(L1 >= 2 ? L1 : 2);
  const double *U2 = c_spher + spher_off(L2 >= 2 ? L2 : 2);
  // pass 1 over d
  const bool t2 = sph && L2 >= 2, t1 = sph && L1 >= 2;
  for (int idx = g; idx < nvec * n1 * s2; idx += G) {
    const int j = idx / (n1 * s2);
    const int rem = idx - j * (n1 * s2);
    const int ds = rem / n1, c = rem - ds * n1;
    double acc;
    if (t2) {
      acc = 0.0;
      for (int d = 0; d < n2; ++d) acc += U2[ds * n2 + d] * v[j * st + c + n1 * d];
    } else {
      acc = v[j * st + c + n1 * ds];
    }
    tmp[j * (n1 * s2) + c + n1 * ds] = acc;
  }
  gsync<G>();
  for (int idx = g; idx < nvec * s1 * s2; idx += G) {
    const int j = idx / (s1 * s2);
    const int rem = idx - j * (s1 * s2);
    const int ds = rem / s1, cs = rem - ds * s1;
    double acc;
    if (t1) {
      acc = 0.0;
      for (int c = 0; c < n1; ++c) acc += U1[cs * n1 + c] * tmp[j * (n1 * s2) + c + n1 * ds];
    } else {
      acc = tmp[j * (n1 * s2) + cs + n1 * ds];
    }
    out[j * ost + cs + s1 * ds] = acc;
  }
  gsync<G>();
}

// triangle position -> (a, b), a >= b, column-major lower triangle
__device__ __forceinline__ void tri_ab(int j, int n, int &a, int &b) {
  int bb = 0, rem = j;
  while (rem >= n - bb) { rem -= n - bb; ++bb; }
  b = bb;
  a = bb + rem;
}

template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(64) eri_class_kernel(EriArgs A) {
  using P = EriPlan<LA, LB, LC, LD>;
  constexpr int G = P::G, QPW = P::QPW;
  const int lane = threadIdx.x;
  const int grp = lane / G, g = lane - grp * G;

  double *base;
  double regbuf[P::MODE == MODE_REG ? P::PER_Q : 1];
  __shared__ double ldsbuf[P::LDS_DOUBLES];
  if constexpr (P::MODE == MODE_REG) base = regbuf;
  else if constexpr (P::MODE == MODE_LDS) base = ldsbuf + grp * P::PER_Q;
  else base = A.gws + (size_t)blockIdx.x * P::PER_Q;

  double *W = base;                    // NR x NC, later VS
  double *TS = base + P::WSZ;          // NR x NCDS
  double *scr = TS + P::TSZ;           // phase scratch

  int m[4], pos[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) { m[k] = (A.perm >> (2 * k)) & 3; pos[m[k]] = k; }
  const long long ntasks = (A.mode == LAUNCH_PRODUCT) ? A.nrows * A.ncols : A.nrows;
  const long long ngroups_total = (ntasks + QPW - 1) / QPW;
  for (long long wg = blockIdx.x; wg < ngroups_total; wg += gridDim.x) {
    const long long tid = wg * QPW + grp;
    const bool valid = tid < ntasks;
    const long long t = valid ? tid : 0;
    long long ir, ic;
    if (A.mode == LAUNCH_PRODUCT) { ic = t / A.nrows; ir = t - ic * A.nrows; }
    else { ir = t; ic = t; }
    const PairDesc RD = A.rdesc[A.rows[ir]];
    const PairDesc CDs = A.cdesc[A.cols[ic]];
    const int caller[4] = {RD.s0, RD.s1, CDs.s0, CDs.s1};
    const DevShell sA = A.shells[caller[m[0]]], sB = A.shells[caller[m[1]]];
    const DevShell sC = A.shells[caller[m[2]]], sD = A.shells[caller[m[3]]];

#pragma unroll
    for (int i = g; i < P::NR * P::NC; i += G) W[i] = 0.0;
    gsync<G>();

    const double Rab2 = (sA.x - sB.x) * (sA.x - sB.x) + (sA.y - sB.y) * (sA.y - sB.y) + (sA.z - sB.z) * (sA.z - sB.z);
    const double Rcd2 = (sC.x - sD.x) * (sC.x - sD.x) + (sC.y - sD.y) * (sC.y - sD.y) + (sC.z - sD.z) * (sC.z - sD.z);
    for (int gd = 0; gd < A.npd; ++gd)
      for (int gc = 0; gc < A.npc; ++gc) {
        const double ec = A.exps[sC.prim0 + gc], ed = A.exps[sD.prim0 + gd];
        const double Q = ec + ed, MuCD = ec * ed / Q;
        double Rq[3] = {(ec * sC.x + ed * sD.x) / Q, (ec * sC.y + ed * sD.y) / Q, (ec * sC.z + ed * sD.z) / Q};
        double Rqc[3] = {Rq[0] - sC.x, Rq[1] - sC.y, Rq[2] - sC.z};
        const double cc = A.cntr[sC.prim0 + gc] * A.cntr[sD.prim0 + gd];
        for (int gb = 0; gb < A.npb; ++gb)
          for (int ga = 0; ga < A.npa; ++ga) {
            const double ea = A.exps[sA.prim0 + ga], eb = A.exps[sB.prim0 + gb];
            const double Pp = ea + eb, MuAB = ea * eb / Pp;
            double Rp[3] = {(ea * sA.x + eb * sB.x) / Pp, (ea * sA.y + eb * sB.y) / Pp, (ea * sA.z + eb * sB.z) / Pp};
            double Rpq[3] = {Rp[0] - Rq[0], Rp[1] - Rq[1], Rp[2] - Rq[2]};
            double Rpa[3] = {Rp[0] - sA.x, Rp[1] - sA.y, Rp[2] - sA.z};
            const double K = exp(-MuAB * Rab2 - MuCD * Rcd2);
            const double PQ = Pp * Q;
            const double den = Pp + Q + A.kappa * PQ;
            const double alpha = PQ / den;
            const double N = 34.986836655249725 / (PQ * sqrt(den));   // 2 pi^(5/2)
            const double e000 = N * A.cntr[sA.prim0 + ga] * A.cntr[sB.prim0 + gb] * cc * K;
            all_blocks<P, 0, 0, G>(W, scr, g, alpha, Rpq, Rpa, Rqc, 0.5 / Pp, 0.5 / Q, e000, A.boys);
          }
      }
    // ket HRR + ket solid harmonics, row chunks of RC
    const double CD[3] = {sC.x - sD.x, sC.y - sD.y, sC.z - sD.z};
    const double AB[3] = {sA.x - sB.x, sA.y - sB.y, sA.z - sB.z};
    const int ncs = nsph(LC, A.spher), nds = nsph(LD, A.spher);
    const int nas = nsph(LA, A.spher), nbs = nsph(LB, A.spher);
    const int ncds = ncs * nds, nabs = nas * nbs;
    for (int r0 = 0; r0 < P::NR; r0 += P::RC) {
      const int rc = imin(P::RC, P::NR - r0);
      double *hout = scr + 2 * imax(P::RC * P::KLEV, P::KC * P::BLEV);   // rc x NCD
      hrr_chunk<LC, LD, G>(W + r0 * P::NC, P::NC, 1, hout, P::NCD, scr, rc, CD, g);
      sph_chunk<LC, LD, G>(hout, P::NCD, TS + r0 * ncds, ncds, scr, rc, A.spher, g);
    }
    // bra HRR + bra solid harmonics, column chunks of KC; VS overlays W
    double *VS = W;
    for (int k0 = 0; k0 < ncds; k0 += P::KC) {
      const int kc = imin(P::KC, ncds - k0);
      double *hout = scr + 2 * imax(P::RC * P::KLEV, P::KC * P::BLEV);   // kc x NAB
      hrr_chunk<LA, LB, G>(TS + k0, 1, ncds, hout, P::NAB, scr, kc, AB, g);
      sph_chunk<LA, LB, G>(hout, P::NAB, VS + k0 * nabs, nabs, scr, kc, A.spher, g);
    }
    // scatter with normalization into the destination block
    if (valid) {
      const int nc_can[4] = {nas, nbs, ncs, nds};
      const int n0 = nc_can[pos[0]], n1 = nc_can[pos[1]], n2 = nc_can[pos[2]], n3 = nc_can[pos[3]];
      const int nrow = (RD.flags & PAIR_SEL) ? RD.nsel : ((RD.flags & PAIR_SAME) ? n0 * (n0 + 1) / 2 : n0 * n1);
      const int ncol = (CDs.flags & PAIR_SEL) ? CDs.nsel : ((CDs.flags & PAIR_SAME) ? n2 * (n2 + 1) / 2 : n2 * n3);
      const double *nA = A.norms + sA.norm0, *nB = A.norms + sB.norm0;
      const double *nC = A.norms + sC.norm0, *nD = A.norms + sD.norm0;
      double *dst = A.out + RD.off + CDs.off;
      const int nwork = (A.mode == LAUNCH_DIAG) ? nrow : nrow * ncol;
      for (int idx = g; idx < nwork; idx += G) {
        int jc, jr;
        if (A.mode == LAUNCH_DIAG) { jr = idx; jc = idx; }
        else { jc = idx / nrow; jr = idx - jc * nrow; }
        int o[4];
        if (RD.flags & PAIR_SEL) {
          const int f = A.sel[RD.sel0 + jr];
          o[1] = f / n0; o[0] = f - o[1] * n0;
        } else if (RD.flags & PAIR_SAME) {
          tri_ab(jr, n0, o[0], o[1]);
        } else {
          o[1] = jr / n0; o[0] = jr - o[1] * n0;
        }
        if (CDs.flags & PAIR_SEL) {
          const int f = A.sel[CDs.sel0 + jc];
          o[3] = f / n2; o[2] = f - o[3] * n2;
        } else if (CDs.flags & PAIR_SAME) {
          tri_ab(jc, n2, o[2], o[3]);
        } else {
          o[3] = jc / n2; o[2] = jc - o[3] * n2;
        }
        const int i0 = o[m[0]], i1 = o[m[1]], i2 = o[m[2]], i3 = o[m[3]];
        const double v = VS[(i0 + nas * i1) + nabs * (i2 + ncs * i3)] * nA[i0] * nB[i1] * nC[i2] * nD[i3];
        if (A.mode == LAUNCH_DIAG) dst[jr] = v;
        else dst[jr + A.ldo * jc] = v;
      }
    }
    gsync<G>();
  }
}

}  // namespace gi
