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
// distinct quartets per tile: their dense R tables within MD4_RBUDGET bytes of LDS
#ifndef MD4_RBUDGET
#define MD4_RBUDGET 9216
#endif
__host__ __device__ constexpr int m4_qmax(int L) {
  return imax(1, imin(4, MD4_RBUDGET / (8 * m4_dsize(L))));
}

#ifndef MD4_ATTR
#define MD4_ATTR
#endif

__device__ __forceinline__ int pick4(const int (&v)[4], int s) {
  return s == 0 ? v[0] : (s == 1 ? v[1] : (s == 2 ? v[2] : v[3]));
}

// a pair record at a uniform address, read into scalar registers
__device__ __forceinline__ QRec load_rec_uniform(const QRec *__restrict__ p) {
  QRec r;
  const int *q = (const int *)p;
  int *o = (int *)&r;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(QRec) / 4); ++i) o[i] = __builtin_amdgcn_readfirstlane(q[i]);
  return r;
}

// 64-bit pointer made provably uniform (scalar registers), so that loads
// through it with a 32-bit lane offset use the scalar-base addressing form
__device__ __forceinline__ const double *uniform_ptr(const double *p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (const double *)(((unsigned long long)hi << 32) | lo);
}
// raw buffer resource over `bytes` bytes from a uniform base (the gfx9
// descriptor word 3, untyped 32-bit data); loads take a per-lane 32-bit offset
// and a uniform (scalar) offset, and return 0 beyond the range
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const double *ubase, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)ubase, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// bra pairs per tile (rows of one column tile handled by one wavefront)
#ifndef MD4_NB
#define MD4_NB 4
#endif
constexpr int M4_NB = MD4_NB;

template <int E, int F>
struct Md4 {
  static constexpr int L = E + F, L1 = L + 1;
  static constexpr int NHE = nherm(E), NHF = nherm(F), NHL = nherm(L);
  static constexpr int NHG = (NHE + 3) / 4, KSF = (NHF + 3) / 4;
  static constexpr int DS = m4_dsize(L), QM = m4_qmax(L), NL1 = L >= 1 ? nherm(L - 1) : 1;
  static constexpr int NRMAX = nab_max(E);
};

// Boys function and top recursion level of one quartet (one lane): the
// scaled F_m (2 pi^(5/2) / (p q sqrt(p + q + kappa p q)) (-2 alpha)^m F_m)
// into F[0..L], P - Q into PQ[0..2]
template <int L>
__device__ __forceinline__ void md4_boys(const PrimRec &PB, const PrimRec &PK, double kappa,
                                         const double *__restrict__ boys, double *F, double *PQ) {
  const double PQx = PB.px - PK.px, PQy = PB.py - PK.py, PQz = PB.pz - PK.pz;
  const double pq = PB.p * PK.p;
  const double den = PB.p + PK.p + kappa * pq;
  const double alpha = pq / den;
  const double N = 34.986836655249725 / (pq * sqrt(den));   // 2 pi^(5/2)
  double Fm[L + 1];
  boys_fm<L>(alpha * (PQx * PQx + PQy * PQy + PQz * PQz), boys, Fm);
  const double m2a = -2.0 * alpha;
  double pw = N;
#pragma unroll
  for (int n = 0; n <= L; ++n) {
    F[n] = Fm[n] * pw;
    pw *= m2a;
  }
  PQ[0] = PQx;
  PQ[1] = PQy;
  PQ[2] = PQz;
}

// PLAIN: every pair of the launch has one primitive pair (uncontracted), so the
// kernel omits the primitive loops
// Waves per SIMD asked of the compiler for the classes whose register count
// lay a few registers above the next occupancy step: it then keeps the MFMA
// accumulators in VGPRs instead of AGPRs and reaches the step with no
// scratch.  Measured per class on the (H2O)16 builds:
//   uncontracted, three waves: (1,6) -21 %, (2,5) -12 %, (2,6) -7 %, (5,2) -6 %,
//     (4,4) -2 %; four waves for (2,3), (3,2), (3,4), (4,0), (4,1) or three for
//     (0,6), (3,5), (5,0) changed nothing or slowed them ((3,4) +15 %);
//   contracted, four waves for (0,3), (1,3), (2,0..2), three for (0,4), (2,3),
//     (2,4), (3,2), (3,4), (4,2), (4,3): 8 to 17 % each, none slower.
// Higher orders spill under either request.
__host__ __device__ constexpr int md4_min_waves(int E, int F, bool plain) {
  if (plain) return ((E == 1 && F == 6) || (E == 2 && F >= 5) || (E == 4 && F == 4) || (E == 5 && F == 2)) ? 3 : 1;
  return ((E == 0 && F == 3) || (E == 1 && F == 3) || (E == 2 && F <= 2)) ? 4
         : ((E == 0 && F == 4) || (E == 2 && (F == 3 || F == 4)) || (E == 3 && (F == 2 || F == 4)) ||
            (E == 4 && (F == 2 || F == 3)))
             ? 3
             : 1;
}

template <int E, int F, bool PLAIN>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(md4_min_waves(E, F, PLAIN)))) MD4_ATTR
md4_kernel(MdArgs A, const ColTile *__restrict__ ctiles, long long nct, const QRec *__restrict__ rrec) {
  using S = Md4<E, F>;
  constexpr int L = S::L, L1 = S::L1, NHE = S::NHE, NHF = S::NHF, NHL = S::NHL, NHG = S::NHG, KSF = S::KSF;
  constexpr int DS = S::DS, QM = S::QM, NL1 = S::NL1;
  __shared__ double s_Rd[QM][DS];              // level-0 R tables, dense
  __shared__ double s_Rp[2][QM][NL1];          // levels >= 1, packed (ping-pong)
  __shared__ double s_F[M4_NB][QM][L + 1];     // scaled Boys values per (bra pair, quartet)
  __shared__ double s_PQ[M4_NB][QM][4];        // P - Q per (bra pair, quartet)
  __shared__ int s_rrec[NHL];
  __shared__ unsigned short s_dix[NHL];        // dense offset of packed index
  __shared__ int s_hoff[NHG * 4];              // dense offset of h (0 for padding)
  __shared__ unsigned s_frow[S::NRMAX + 4];    // byte offset of each bra row's full-block index

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
  // the last h step pads h beyond NHE (rows h = 4 hg + lq of the stage-2 A operand)
  const bool hlast_ok = 4 * (NHG - 1) + lq < NHE;
  constexpr bool hg_last_mask = (NHE % 4) != 0;

  // XCD-aware order: consecutive tiles (one row block) on one XCD's L2
  const int G = gridDim.x;
  const int w = blockIdx.x;
  const long long v0 = (G % 8 == 0) ? (long long)(w % 8) * (G / 8) + w / 8 : w;
  const long long rs = A.trans ? A.ldo : 1, cs = A.trans ? 1 : A.ldo;
  // row blocks of A.kt (1 .. M4_NB) bra pairs: fewer per tile where a launch
  // has too few tiles to fill the device
  const int nbr = A.kt;
  const long long nrb = (A.nrows + nbr - 1) / nbr;
  const long long ntiles = nrb * nct;

  // R tables of the nq quartets of bra pair b (Boys values in s_F[b], s_PQ[b]):
  // the recursion R^(n)_{tuv} = coef R^(n+1)_{i2} + PQ_dir R^(n+1)_{i1},
  // levels L-1 .. 0, lpq lanes per quartet; level 0 into the dense tables
  auto build_r = [&](int b, int nq, unsigned actmask) {
    const int lpq = nq <= 1 ? 64 : (nq == 2 ? 32 : 16);
    const int ql = lane / lpq, sub = lane - ql * lpq;
    const bool act = ql < nq && ((actmask >> ql) & 1u);
    __syncthreads();   // Boys values written; earlier readers of the tables are done
    if (act && L > 0 && sub == 0) s_Rp[L % 2][ql][0] = s_F[b][ql][L];
    if (act && L == 0 && sub == 0) s_Rd[ql][0] = s_F[b][ql][0];
#pragma unroll
    for (int n = L - 1; n >= 0; --n) {
      __syncthreads();
      const int cnt = nherm(L - n);
      if (act) {
        const double *prv = s_Rp[(n + 1) % 2][ql];
        for (int idx = sub; idx < cnt; idx += lpq) {
          double r;
          if (idx == 0) {
            r = s_F[b][ql][n];
          } else {
            const int rec = s_rrec[idx];
            const int i1 = rec & 0x7ff, i2 = (rec >> 11) & 0x7ff, dir = (rec >> 22) & 3, coef = rec >> 24;
            r = (double)coef * prv[i2] + s_PQ[b][ql][dir] * prv[i1];
          }
          if (n == 0)
            s_Rd[ql][s_dix[idx]] = r;
          else
            s_Rp[n % 2][ql][idx] = r;
        }
      }
    }
    __syncthreads();
  };

  for (long long t = v0; t < ntiles; t += G) {
    const long long rb = t / nct, ct = t - rb * nct;
    const long long ir0 = rb * nbr;
    const int nb = (int)((A.nrows - ir0) < nbr ? (A.nrows - ir0) : nbr);
    const ColTile T = ctiles[ct];
    const int nq = T.nq;
    // this lane's slot: ket pair, column (shared by the tile's bra pairs)
    const int myrt = pick4(T.rt, lb);
    const int myc0 = pick4(T.c0, lb);
    const QRec CD = A.crec[pick4(T.jq, myrt)];
    const int c = myc0 + li;
    const bool cval = myc0 >= 0 && c < CD.nrow;
    const int fc = cval ? pair_full(CD, A.sel, c) : 0;
    const int ldk = CD.n0 * CD.n1;

    // stage-1 B operand of ket primitive kp: (-1)^{|k|} ET_ket(k, c).  Every
    // lane loads (index 0 where k or c is outside the table) and scales by
    // +-1 or 0 afterwards: the loads are all in flight together instead of
    // each waiting behind a branch
    double eb[KSF];
    auto load_eb = [&](int kp) {
      const bool kval = cval && kp < CD.nprim;
      const int kpc = kp < CD.nprim ? kp : CD.nprim - 1;
      const PrimRec PKm = kpc == 0 ? first_prim(CD) : A.prim[CD.prim0 + kpc];
      const double *etk = A.et + PKm.et;
#pragma unroll
      for (int kg = 0; kg < KSF; ++kg) {
        const int k = 4 * kg + lq;
        const bool ok = kval && k < NHF;
        eb[kg] = etk[ok ? k * ldk + fc : 0];
      }
#pragma unroll
      for (int kg = 0; kg < KSF; ++kg) {
        const bool ok = kval && 4 * kg + lq < NHF;
        eb[kg] *= ok ? (((ksgn >> kg) & 1u) ? -1.0 : 1.0) : 0.0;
      }
    };
    // stage 1: X(h, c) += sum_k R_{h+k} B(k, c) from the dense table of this
    // lane's slot; k step kg updates NHG independent accumulators (the step's
    // offset passes through an empty asm statement so that the NHG addresses
    // are formed in the step instead of all NHG x KSF hoisted into registers)
    double X[NHG];
    auto stage1 = [&]() {
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
    };
    // bra rows of RB: byte offset of each row's full-block index
    auto fill_rows = [&](const QRec &RB) {
      for (int r = lane; r < RB.nrow; r += 64) s_frow[r] = 8u * (unsigned)pair_full(RB, A.sel, r);
      if (lane < 4) s_frow[RB.nrow + lane] = 0u;   // padding rows of the last row group
      __syncthreads();
    };
    // stage 2 for one bra primitive pair: V(r, c) = sum_h ET_bra(h, r) X(h, c),
    // four bra rows per step (the A operand's row 4 rg + li); the table of h step
    // hg starts at a uniform address, each lane reads at a 32-bit offset.
    // s2_open + load_rows(0) issue the first row group's loads as soon as the
    // bra rows are known, so that they complete under the R recursion and
    // stage 1; each row group then loads the next one while it multiplies.
    __amdgpu_buffer_rsrc_t s2_ub = make_rsrc(uniform_ptr(A.et), 0u);
    unsigned s2_hstep = 0, s2_lrow = 0, s2_lrow_last = 0;
    double ea[NHG];
    auto s2_open = [&](const QRec &RB, const double *etb) {
      const int ldb = RB.n0 * RB.n1;
      s2_ub = make_rsrc(uniform_ptr(etb), 8u * (unsigned)(NHE * ldb));
      s2_hstep = 32u * (unsigned)ldb;   // 4 h rows, bytes
      s2_lrow = 8u * (unsigned)(lq * ldb);
      // the last h step reads row min(h, NHE - 1) (inside the block), masked
      s2_lrow_last = 8u * (unsigned)((imin(4 * (NHG - 1) + lq, NHE - 1) - 4 * (NHG - 1)) * ldb);
    };
    auto load_rows = [&](int rg) {
      const unsigned fro = s_frow[4 * rg + li];
#pragma unroll
      for (int hg = 0; hg < NHG; ++hg) {
        const unsigned o = (hg == NHG - 1 && (NHE % 4) != 0) ? s2_lrow_last + fro : s2_lrow + fro;
        ea[hg] = buf_ld(s2_ub, o, (unsigned)hg * s2_hstep);
      }
    };
    // two accumulator chains (even and odd h steps)
    auto stage2 = [&](const QRec &RB, bool accumulate) {
      const int nr = RB.nrow;
      double *const out = A.out + RB.off + CD.off;
      const int nrg = (nr + 3) / 4;
      for (int rg = 0; rg < nrg; ++rg) {
        double e[NHG];
#pragma unroll
        for (int hg = 0; hg < NHG; ++hg) e[hg] = ea[hg];
        if (hg_last_mask) e[NHG - 1] = hlast_ok ? e[NHG - 1] : 0.0;
        if (rg + 1 < nrg) load_rows(rg + 1);
        double v0 = 0.0, v1 = 0.0;
#pragma unroll
        for (int hg = 0; hg < NHG; ++hg) {
          if (hg & 1)
            v1 = __builtin_amdgcn_mfma_f64_4x4x4f64(e[hg], X[hg], v1, 0, 0, 0);
          else
            v0 = __builtin_amdgcn_mfma_f64_4x4x4f64(e[hg], X[hg], v0, 0, 0, 0);
        }
        const int ro = 4 * rg + lq;
        if (cval && ro < nr) {
          double *d = out + rs * ro + cs * c;
          *d = accumulate ? *d + (v0 + v1) : v0 + v1;
        }
      }
    };

    if constexpr (PLAIN) {
      // uncontracted tile: one Boys pass for every (bra pair, ket pair)
      // quartet of the tile, then per bra pair the R tables and both stages
      load_eb(0);
      __syncthreads();   // earlier readers of s_F / s_PQ are done
      if (lane < nb * nq) {
        const int b = lane / nq, q = lane - b * nq;
        const QRec RB = A.rrec[ir0 + b];
        const QRec CQ = A.crec[pick4(T.jq, q)];
        md4_boys<L>(first_prim(RB), first_prim(CQ), A.kappa, A.boys, s_F[b][q], s_PQ[b][q]);
      }
      for (int b = 0; b < nb; ++b) {
        const QRec RB = load_rec_uniform(rrec + ir0 + b);
        fill_rows(RB);
        s2_open(RB, A.et + RB.et);
        load_rows(0);
        build_r(b, nq, 0xfu);
#pragma unroll
        for (int hg = 0; hg < NHG; ++hg) X[hg] = 0.0;
        stage1();
        stage2(RB, false);
      }
    } else {
      // contracted pairs: per bra pair and bra primitive pair, X summed over
      // the ket primitive pairs; the block accumulates over bra primitives.
      // The Boys values of up to M4_NB ket primitive steps are evaluated in
      // one pass (one lane per step and quartet) before their R tables.
      int maxnp = 1;
      for (int qq = 0; qq < nq; ++qq) maxnp = imax(maxnp, A.crec[pick4(T.jq, qq)].nprim);
      for (int b = 0; b < nb; ++b) {
        const QRec RB = load_rec_uniform(rrec + ir0 + b);
        fill_rows(RB);
        for (int bp = 0; bp < RB.nprim; ++bp) {
          const PrimRec PB = bp == 0 ? first_prim(RB) : A.prim[RB.prim0 + bp];
          s2_open(RB, A.et + PB.et);
          load_rows(0);
#pragma unroll
          for (int hg = 0; hg < NHG; ++hg) X[hg] = 0.0;
          for (int kp0 = 0; kp0 < maxnp; kp0 += M4_NB) {
            const int nkk = imin(M4_NB, maxnp - kp0);
            __syncthreads();   // earlier readers of the R tables and s_F are done
            if (lane < nkk * nq) {
              const int kk = lane / nq, q = lane - kk * nq, kp = kp0 + kk;
              const QRec CQ = A.crec[pick4(T.jq, q)];
              if (kp < CQ.nprim) {
                const PrimRec PK = kp == 0 ? first_prim(CQ) : A.prim[CQ.prim0 + kp];
                md4_boys<L>(PB, PK, A.kappa, A.boys, s_F[kk][q], s_PQ[kk][q]);
              }
            }
            for (int kk = 0; kk < nkk; ++kk) {
              const int kp = kp0 + kk;
              unsigned act = 0;
              for (int qq = 0; qq < nq; ++qq)
                if (kp < A.crec[pick4(T.jq, qq)].nprim) act |= 1u << qq;
              build_r(kk, nq, act);
              load_eb(kp);
              stage1();
            }
          }
          stage2(RB, bp > 0);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Thin-ket variant: classes whose ket pairs are s-s pairs (F = 0, one Hermite
// function and one column per ket pair).  Stage 1 reduces to
// X(h, c) = R_h(quartet of c) ET_ket(c), a scale instead of a product, so the
// four blocks of an instruction need not share a quartet: a tile is up to
// M4_NB bra pairs x M4S_NK consecutive ket pairs, every column its own quartet, and
// stage 2 (the bra table, shared by all blocks, times X) fills all sixteen
// columns of the instruction.  The R tables are packed (index h, no h + k
// addressing), sixteen per bra pair, four lanes per quartet; even recursion
// levels in one table, odd levels in the other, level 0 last.  The Boys
// values of the tile's quartets stay in the registers of the lane that
// evaluated them and reach the recursion lanes by cross-lane reads.
constexpr int M4S_NK = 16;

template <int E, bool PLAIN>
__global__ void __launch_bounds__(64) MD4_ATTR md4s_kernel(MdArgs A, const QRec *__restrict__ rrec) {
  constexpr int L = E, NHE = nherm(E), NHG = (NHE + 3) / 4;
  constexpr int NL1 = L >= 1 ? nherm(L - 1) : 1;
  constexpr int SB = NHE | 1, SA = NL1 | 1;   // odd row strides (LDS banks)
  constexpr int NRMAX = nab_max(E);
  __shared__ double s_Rb[M4S_NK * SB];        // even levels, level 0 last
  __shared__ double s_Ra[M4S_NK * SA];        // odd levels
  __shared__ double s_pq[M4S_NK][4];          // P - Q of the bra pair's quartets
  __shared__ int s_rrec[NHE];
  __shared__ unsigned s_frow[NRMAX + 4];

  const int lane = threadIdx.x;
  const int lq = lane >> 4, li = lane & 3, col = lane & 15;
  for (int e = lane; e < NHE; e += 64) s_rrec[e] = c_rrec[e];
  __syncthreads();
  const bool hlast_ok = 4 * (NHG - 1) + lq < NHE;
  constexpr bool hg_last_mask = (NHE % 4) != 0;

  const int G = gridDim.x;
  const int w = blockIdx.x;
  const long long v0 = (G % 8 == 0) ? (long long)(w % 8) * (G / 8) + w / 8 : w;
  const long long rs = A.trans ? A.ldo : 1;
  const int nbr = A.kt;   // bra pairs per row block, 1 .. M4_NB
  const long long nrb = (A.nrows + nbr - 1) / nbr;
  const long long nct = (A.ncols + M4S_NK - 1) / M4S_NK;
  const long long ntiles = nrb * nct;

  // the Boys values (scaled F_0 .. F_L) and P - Q of this lane's quartet
  double Fq[L + 1], PQq[3];
  auto clear_boys = [&]() {
#pragma unroll
    for (int n = 0; n <= L; ++n) Fq[n] = 0.0;
    PQq[0] = PQq[1] = PQq[2] = 0.0;
  };
  // R tables of the sixteen quartets whose Boys values sit in lanes
  // src0 .. src0 + 15: lanes 4 j .. 4 j + 3 recurse quartet j
  auto build_r = [&](int src0) {
    const int j = lane >> 2, sub = lane & 3;
    const int src = src0 + j;
    const double px_ = __shfl(PQq[0], src), py_ = __shfl(PQq[1], src), pz_ = __shfl(PQq[2], src);
    const double ftop = __shfl(Fq[L], src);
    __syncthreads();   // earlier readers of the tables are done
    if (sub == 0) {
      s_pq[j][0] = px_;
      s_pq[j][1] = py_;
      s_pq[j][2] = pz_;
      ((L & 1) ? s_Ra + j * SA : s_Rb + j * SB)[0] = ftop;
    }
#pragma unroll
    for (int n = L - 1; n >= 0; --n) {
      const double fn = __shfl(Fq[n], src);
      __syncthreads();
      const int cnt = nherm(L - n);
      const double *prv = ((n + 1) & 1) ? s_Ra + j * SA : s_Rb + j * SB;
      double *cur = (n & 1) ? s_Ra + j * SA : s_Rb + j * SB;
      for (int idx = sub; idx < cnt; idx += 4) {
        double r;
        if (idx == 0) {
          r = fn;
        } else {
          const int rec = s_rrec[idx];
          const int i1 = rec & 0x7ff, i2 = (rec >> 11) & 0x7ff, dir = (rec >> 22) & 3, coef = rec >> 24;
          r = (double)coef * prv[i2] + s_pq[j][dir] * prv[i1];
        }
        cur[idx] = r;
      }
    }
    __syncthreads();
  };
  // stage 1: X(h, c) (+)= R_h ET_ket(c) for this lane's h = 4 hg + lq, c = col
  double X[NHG];
  auto stage1 = [&](double etv, bool add) {
    const double *Rc = s_Rb + col * SB;
#pragma unroll
    for (int hg = 0; hg < NHG; ++hg) {
      const int h = 4 * hg + lq;
      const double v = Rc[h < NHE ? h : NHE - 1] * etv;
      const double xv = (hg == NHG - 1 && hg_last_mask && !hlast_ok) ? 0.0 : v;
      X[hg] = add ? X[hg] + xv : xv;
    }
  };
  auto fill_rows = [&](const QRec &RB) {
    for (int r = lane; r < RB.nrow; r += 64) s_frow[r] = 8u * (unsigned)pair_full(RB, A.sel, r);
    if (lane < 4) s_frow[RB.nrow + lane] = 0u;
    __syncthreads();
  };
  // stage 2 as in md4_kernel (first row group loaded ahead): V(r, c) =
  // sum_h ET_bra(h, r) X(h, c), column c of this lane the single column of
  // its ket pair
  __amdgpu_buffer_rsrc_t s2_ub = make_rsrc(uniform_ptr(A.et), 0u);
  unsigned s2_hstep = 0, s2_lrow = 0, s2_lrow_last = 0;
  double ea[NHG];
  auto s2_open = [&](const QRec &RB, const double *etb) {
    const int ldb = RB.n0 * RB.n1;
    s2_ub = make_rsrc(uniform_ptr(etb), 8u * (unsigned)(NHE * ldb));
    s2_hstep = 32u * (unsigned)ldb;
    s2_lrow = 8u * (unsigned)(lq * ldb);
    s2_lrow_last = 8u * (unsigned)((imin(4 * (NHG - 1) + lq, NHE - 1) - 4 * (NHG - 1)) * ldb);
  };
  auto load_rows = [&](int rg) {
    const unsigned fro = s_frow[4 * rg + li];
#pragma unroll
    for (int hg = 0; hg < NHG; ++hg) {
      const unsigned o = (hg == NHG - 1 && (NHE % 4) != 0) ? s2_lrow_last + fro : s2_lrow + fro;
      ea[hg] = buf_ld(s2_ub, o, (unsigned)hg * s2_hstep);
    }
  };
  auto stage2 = [&](const QRec &RB, double *outc, bool cval, bool accumulate) {
    const int nr = RB.nrow;
    double *const out = outc + RB.off;
    const int nrg = (nr + 3) / 4;
    for (int rg = 0; rg < nrg; ++rg) {
      double e[NHG];
#pragma unroll
      for (int hg = 0; hg < NHG; ++hg) e[hg] = ea[hg];
      if (hg_last_mask) e[NHG - 1] = hlast_ok ? e[NHG - 1] : 0.0;
      if (rg + 1 < nrg) load_rows(rg + 1);
      double v0 = 0.0, v1 = 0.0;
#pragma unroll
      for (int hg = 0; hg < NHG; ++hg) {
        if (hg & 1)
          v1 = __builtin_amdgcn_mfma_f64_4x4x4f64(e[hg], X[hg], v1, 0, 0, 0);
        else
          v0 = __builtin_amdgcn_mfma_f64_4x4x4f64(e[hg], X[hg], v0, 0, 0, 0);
      }
      const int ro = 4 * rg + lq;
      if (cval && ro < nr) {
        double *d = out + rs * ro;
        *d = accumulate ? *d + (v0 + v1) : v0 + v1;
      }
    }
  };

  for (long long t = v0; t < ntiles; t += G) {
    const long long rb = t / nct, ct = t - rb * nct;
    const long long ir0 = rb * nbr, jc0 = ct * M4S_NK;
    const int nb = (int)((A.nrows - ir0) < nbr ? (A.nrows - ir0) : nbr);
    const int nk = (int)((A.ncols - jc0) < M4S_NK ? (A.ncols - jc0) : M4S_NK);
    // this lane's column: ket pair jc0 + col, its single orbital pair
    const bool cval = col < nk;
    const QRec CD = A.crec[jc0 + (cval ? col : 0)];
    const int fc = cval ? pair_full(CD, A.sel, 0) : 0;
    double *const outc = A.out + CD.off;

    if constexpr (PLAIN) {
      // one Boys pass for every quartet of the tile: lane 16 b + j takes
      // bra pair b, ket pair j
      {
        const int b = lane >> 4, j = lane & 15;
        clear_boys();
        if (b < nb && j < nk)
          md4_boys<L>(first_prim(A.rrec[ir0 + b]), first_prim(A.crec[jc0 + j]), A.kappa, A.boys, Fq, PQq);
      }
      const double etv = cval ? A.et[CD.et + fc] : 0.0;
      for (int b = 0; b < nb; ++b) {
        const QRec RB = load_rec_uniform(rrec + ir0 + b);
        fill_rows(RB);
        s2_open(RB, A.et + RB.et);
        load_rows(0);
        build_r(16 * b);
        stage1(etv, false);
        stage2(RB, outc, cval, false);
      }
    } else {
      // contracted pairs: per bra pair and bra primitive pair, X summed over
      // the ket primitive pairs (lanes 0 .. 15 evaluate the Boys values of the
      // sixteen quartets of one primitive step)
      int maxnp = 1;
      for (int j = 0; j < nk; ++j) maxnp = imax(maxnp, A.crec[jc0 + j].nprim);
      for (int b = 0; b < nb; ++b) {
        const QRec RB = load_rec_uniform(rrec + ir0 + b);
        fill_rows(RB);
        for (int bp = 0; bp < RB.nprim; ++bp) {
          const PrimRec PB = bp == 0 ? first_prim(RB) : A.prim[RB.prim0 + bp];
          s2_open(RB, A.et + PB.et);
          load_rows(0);
          for (int kp = 0; kp < maxnp; ++kp) {
            clear_boys();
            if (lane < nk) {
              const QRec CQ = A.crec[jc0 + lane];
              if (kp < CQ.nprim) {
                const PrimRec PK = kp == 0 ? first_prim(CQ) : A.prim[CQ.prim0 + kp];
                md4_boys<L>(PB, PK, A.kappa, A.boys, Fq, PQq);
              }
            }
            build_r(0);
            double etv = 0.0;
            if (cval && kp < CD.nprim) {
              const PrimRec PK = kp == 0 ? first_prim(CD) : A.prim[CD.prim0 + kp];
              etv = A.et[PK.et + fc];
            }
            stage1(etv, kp > 0);
          }
          stage2(RB, outc, cval, bp > 0);
        }
      }
    }
  }
}

}  // namespace gi
