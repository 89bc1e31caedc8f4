// ERI launch layer of the Hermite-matrix engine: deduplicates the shell pairs
// of a request into a pair table, builds their Hermite coefficient tables
// (k_build_et), bins bra pairs by E = l0 + l1 and ket pairs by F = l2 + l3,
// and launches md_kernel<E, F> on the Cartesian product (or the element-wise
// pairing) of every bra bin with every ket bin.  One instantiation per
// (E, F), listed by gigpu/eri_gen.py in md_classes.inc.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <unordered_map>
#include "gi_context.h"
#include "md_generic.h"
#include "md4_kernel.h"
#include "grad_kernels.h"
#include "onee_kernels.h"
#include "thc_kernels.h"

namespace gi {

__constant__ double c_spher[SPHER_TOTAL];
__constant__ int c_herm[NHRMAX];
__constant__ int c_rrec[NHRMAX];

// ------------------------------------------------------------- pair tables
// ET(h, ma + n0 mb) of one primitive pair: two-centre 1D Hermite coefficients
//   E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}
//   E^{i,j+1}_t = E^{ij}_{t-1}/(2p) + X_PB E^{ij}_t + (t+1) E^{ij}_{t+1}
// multiplied over x, y, z, transformed to real solid harmonics on both
// shells (unit-sphere normalized, m = -l..l, auto2e_SpherAO.F90 convention),
// times the AO normalization factors, the contraction coefficients and
// exp(-ab/p |AB|^2).  One 64-lane workgroup per primitive pair.
constexpr int ET_HCH = 8;   // h values per LDS chunk
__global__ void __launch_bounds__(64) k_build_et(const DevShell *__restrict__ shells, const double *__restrict__ exps,
                                                 const double *__restrict__ cntr, const double *__restrict__ norms,
                                                 int spher, const EtJob *__restrict__ jobs, PrimRec *prims,
                                                 double *et) {
  constexpr int N1D = (MAXL + 1) * (MAXL + 1) * (2 * MAXL + 1);
  __shared__ double s_e1[3][N1D];
  __shared__ double s_t[ncart(MAXL) * ncart(MAXL) * ET_HCH];
  const int lane = threadIdx.x;
  const EtJob J = jobs[blockIdx.x];
  const DevShell A = shells[J.s0], B = shells[J.s1];
  const int la = A.l, lb = B.l, lab = la + lb;
  const double a = exps[A.prim0 + J.ia], b = exps[B.prim0 + J.ib];
  const double p = a + b;
  const double P[3] = {(a * A.x + b * B.x) / p, (a * A.y + b * B.y) / p, (a * A.z + b * B.z) / p};
  const double AB2 = (A.x - B.x) * (A.x - B.x) + (A.y - B.y) * (A.y - B.y) + (A.z - B.z) * (A.z - B.z);
  const double pref = cntr[A.prim0 + J.ia] * cntr[B.prim0 + J.ib] * exp(-a * b / p * AB2);
  if (lane == 0) prims[J.prim] = PrimRec{P[0], P[1], P[2], p, J.et, 0};
  const int s_i = (lb + 1) * (lab + 1), s_j = lab + 1;   // E1(i, j, t) at i*s_i + j*s_j + t
  if (lane < 3) {
    const double Ac = lane == 0 ? A.x : (lane == 1 ? A.y : A.z);
    const double Bc = lane == 0 ? B.x : (lane == 1 ? B.y : B.z);
    const double xpa = P[lane] - Ac, xpb = P[lane] - Bc, i2p = 0.5 / p;
    double *E1 = s_e1[lane];
    for (int k = 0; k < (la + 1) * s_i; ++k) E1[k] = 0.0;
    E1[0] = 1.0;
    for (int i = 0; i < la; ++i)
      for (int t = 0; t <= i + 1; ++t) {
        double v = xpa * (t <= i ? E1[i * s_i + t] : 0.0);
        if (t >= 1) v += i2p * E1[i * s_i + t - 1];
        if (t + 1 <= i) v += (double)(t + 1) * E1[i * s_i + t + 1];
        E1[(i + 1) * s_i + t] = v;
      }
    for (int i = 0; i <= la; ++i)
      for (int j = 0; j < lb; ++j)
        for (int t = 0; t <= i + j + 1; ++t) {
          const int o = i * s_i + j * s_j;
          double v = xpb * (t <= i + j ? E1[o + t] : 0.0);
          if (t >= 1) v += i2p * E1[o + t - 1];
          if (t + 1 <= i + j) v += (double)(t + 1) * E1[o + t + 1];
          E1[o + s_j + t] = v;
        }
  }
  __syncthreads();
  const int nc0 = ncart(la), nc1 = ncart(lb);
  const bool t0 = spher && la >= 2, t1 = spher && lb >= 2;
  const int n0 = t0 ? 2 * la + 1 : nc0, n1 = t1 ? 2 * lb + 1 : nc1;
  const double *U0 = c_spher + spher_off(la >= 2 ? la : 2);
  const double *U1 = c_spher + spher_off(lb >= 2 ? lb : 2);
  const double *nA = norms + A.norm0, *nB = norms + B.norm0;
  const int NH = nherm(lab);
  const int nab = n0 * n1;
  double *out = et + J.et;
  for (int h0 = 0; h0 < NH; h0 += ET_HCH) {
    // pass 1: T(a, mb, hh) = sum_b U1(mb, b) E(a, b, h0 + hh)
    for (int e = lane; e < nc0 * n1 * ET_HCH; e += 64) {
      const int hh = e % ET_HCH, rest = e / ET_HCH;
      const int mb = rest % n1, ia = rest / n1;
      const int h = h0 + hh;
      double acc = 0.0;
      if (h < NH) {
        const int th = c_herm[h];
        const int tt = px(th), uu = py(th), vv = pz(th);
        const int ca = cart_of(la, ia);
        const int ax = px(ca), ay = py(ca), az = pz(ca);
        for (int ib = 0; ib < nc1; ++ib) {
          const double u = t1 ? U1[mb * nc1 + ib] : (ib == mb ? 1.0 : 0.0);
          if (u == 0.0) continue;
          const int cb = cart_of(lb, ib);
          const int bx = px(cb), by = py(cb), bz = pz(cb);
          if (tt > ax + bx || uu > ay + by || vv > az + bz) continue;
          acc += u * s_e1[0][ax * s_i + bx * s_j + tt] * s_e1[1][ay * s_i + by * s_j + uu] *
                 s_e1[2][az * s_i + bz * s_j + vv];
        }
      }
      s_t[e] = acc;
    }
    __syncthreads();
    // pass 2: ET(h, ma + n0 mb) = pref nA(ma) nB(mb) sum_a U0(ma, a) T(a, mb, hh)
    for (int e = lane; e < nab * ET_HCH; e += 64) {
      const int hh = e / nab, r = e - hh * nab;
      const int h = h0 + hh;
      if (h >= NH) continue;
      const int mb = r / n0, ma = r - mb * n0;
      double acc = 0.0;
      if (t0) {
        for (int ia = 0; ia < nc0; ++ia) acc += U0[ma * nc0 + ia] * s_t[(ia * n1 + mb) * ET_HCH + hh];
      } else {
        acc = s_t[(ma * n1 + mb) * ET_HCH + hh];
      }
      out[(long long)h * nab + r] = acc * pref * nA[ma] * nB[mb];
    }
    __syncthreads();
  }
}


// packed record of entry i of a launch index list (side 0: bra list, 1: ket list)
__global__ void k_pack_recs(const PairDesc *__restrict__ rd, const PairDesc *__restrict__ cd,
                            const int *__restrict__ idx, const unsigned char *__restrict__ side, long long n,
                            const PairHdr *__restrict__ hdr, const PrimRec *__restrict__ prim, QRec *out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PairDesc d = side[i] ? cd[idx[i]] : rd[idx[i]];
  const PairHdr h = hdr[d.pt];
  const PrimRec p = prim[h.prim0];
  const int nrow = (d.flags & PAIR_SEL) ? d.nsel : ((d.flags & PAIR_SAME) ? h.n0 * (h.n0 + 1) / 2 : h.n0 * h.n1);
  out[i] = QRec{p.px, p.py, p.pz, p.p, p.et, d.off, nrow, d.flags, d.sel0, h.n0, h.n1, h.nprim, h.prim0, 0};
}

using LaunchFn = void (*)(const MdArgs &, int grid, hipStream_t);

template <int E, int F>
static void launch_md(const MdArgs &a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((md_kernel<E, F>), dim3((unsigned)grid), dim3(64), 0, s, a);
}

template <int E, int F>
static void launch_lane(const MdArgs &a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((md_lane_kernel<E, F>), dim3((unsigned)grid), dim3(256), 0, s, a);
}

using Launch4Fn = void (*)(const MdArgs &, const ColTile *, long long, bool plain, int grid, hipStream_t);

template <int E, int F>
static void launch_md4(const MdArgs &a, const ColTile *ct, long long nct, bool plain, int grid, hipStream_t s) {
  if (plain)
    hipLaunchKernelGGL((md4_kernel<E, F, true>), dim3((unsigned)grid), dim3(64), 0, s, a, ct, nct, a.rrec);
  else
    hipLaunchKernelGGL((md4_kernel<E, F, false>), dim3((unsigned)grid), dim3(64), 0, s, a, ct, nct, a.rrec);
}

using Launch4sFn = void (*)(const MdArgs &, bool plain, int grid, hipStream_t);

template <int E>
static void launch_md4s(const MdArgs &a, bool plain, int grid, hipStream_t s) {
  if (plain)
    hipLaunchKernelGGL((md4s_kernel<E, true>), dim3((unsigned)grid), dim3(64), 0, s, a, a.rrec);
  else
    hipLaunchKernelGGL((md4s_kernel<E, false>), dim3((unsigned)grid), dim3(64), 0, s, a, a.rrec);
}
static const Launch4sFn md4s_table[MD_EMAX + 1] = {&launch_md4s<0>, &launch_md4s<1>, &launch_md4s<2>, &launch_md4s<3>,
                                                   &launch_md4s<4>, &launch_md4s<5>, &launch_md4s<6>};

static LaunchFn md_table[MD_EMAX + 1][MD_EMAX + 1];
static LaunchFn lane_table[MD_EMAX + 1][MD_EMAX + 1];   // null where a class has no lane kernel
static Launch4Fn md4_table[MD_EMAX + 1][MD_EMAX + 1];

static void init_md_table() {
  static bool done = false;
  if (done) return;
#define GI_MD(e, f) md_table[e][f] = &launch_md<e, f>; md4_table[e][f] = &launch_md4<e, f>;
#define GI_LANE(e, f) lane_table[e][f] = &launch_lane<e, f>;
#include "md_classes.inc"
#undef GI_MD
#undef GI_LANE
  done = true;
}

// one (bra pair, ket pair) quartet per workgroup of md_generic_kernel
static void launch_generic(int E, int F, MdArgs a, hipStream_t s) {
  if (E > GEN_EMAX || F > GEN_EMAX || E + F > BOYS_MAXM)
    throw Error{60, "no ERI kernel for Hermite orders " + std::to_string(E) + ", " + std::to_string(F)};
  a.kt = 1;
  a.trans = 0;
  a.ntiles = (a.mode == LAUNCH_PRODUCT) ? a.nrows * a.ncols : a.nrows;
  const long long grid = std::max<long long>(1, std::min<long long>(a.ntiles, 16384));
  hipLaunchKernelGGL(md_generic_kernel, dim3((unsigned)grid), dim3(GEN_NT), 0, s, a, E, F);
}

void launch_md_class(int E, int F, const MdArgs &a, int grid, hipStream_t s) {
  init_md_table();
  if (E < 0 || F < 0) throw Error{60, "no ERI kernel for negative Hermite orders"};
  if (E > MD_EMAX || F > MD_EMAX) {
    launch_generic(E, F, a, s);
    return;
  }
  if (!md_table[E][F])
    throw Error{60, "no ERI kernel for Hermite orders " + std::to_string(E) + ", " + std::to_string(F)};
  md_table[E][F](a, grid, s);
}

void upload_constants(Context &C) {
  init_md_table();
  std::vector<double> U(SPHER_TOTAL, 0.0);
  for (int l = 2; l <= MAXL; ++l) spher_u_host(l, U.data() + spher_off(l));
  GI_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_spher), U.data(), sizeof(double) * SPHER_TOTAL));
  // Hermite index tables
  std::vector<int> herm(NHRMAX), rrec(NHRMAX, 0);
  for (int i = 0; i < NHRMAX; ++i) {
    const int p = herm_of(i);
    herm[i] = p;
    const int t = px(p), u = py(p), v = pz(p);
    if (i == 0) continue;
    int dir, c;
    if (t > 0) { dir = 0; c = t; }
    else if (u > 0) { dir = 1; c = u; }
    else { dir = 2; c = v; }
    const int i1 = hidx(t - (dir == 0), u - (dir == 1), v - (dir == 2));
    const int i2 = (c >= 2) ? hidx(t - 2 * (dir == 0), u - 2 * (dir == 1), v - 2 * (dir == 2)) : 0;
    rrec[i] = i1 | (i2 << 11) | (dir << 22) | ((c - 1) << 24);
  }
  GI_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_herm), herm.data(), sizeof(int) * NHRMAX));
  GI_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_rrec), rrec.data(), sizeof(int) * NHRMAX));
  std::vector<double> table;
  boys_table_host(BOYS_MAXM, table);
  C.dboys.upload(table, C.stream);
  GI_HIP(hipStreamSynchronize(C.stream));
}

// FP64 operations of the R table of one primitive quartet of total order L:
// Boys interpolation and recursion + 3 per recurrence entry
static double rtable_flops(int L) {
  double n = 0;
  for (int m = 0; m <= L; ++m) n += nherm(m);
  return 14.0 + 3.0 * L + 3.0 * n;
}

static int pair_rows_host(const PairDesc &d, int n0, int n1) {
  if (d.flags & PAIR_SEL) return d.nsel;
  if (d.flags & PAIR_SAME) return n0 * (n0 + 1) / 2;
  return n0 * n1;
}

// Column tiles of an md4 launch: the ket pairs' columns in groups of four,
// four groups per tile, at most qmax distinct ket pairs per tile.  Returns the
// number of R tables the tiles need (sum of their distinct ket pairs).
static long long build_col_tiles(const std::vector<int> &ncol, int qmax, std::vector<ColTile> &out) {
  ColTile cur{};
  int nslot = 0;
  long long nrt = 0;
  auto reset = [&]() {
    cur = ColTile{};
    for (int s = 0; s < 4; ++s) {
      cur.c0[s] = -1;
      cur.jq[s] = 0;
    }
    cur.nq = 0;
    nslot = 0;
  };
  auto flush = [&]() {
    if (nslot == 0) return;
    for (int q = cur.nq; q < 4; ++q) cur.jq[q] = cur.jq[0];
    nrt += cur.nq;
    out.push_back(cur);
    reset();
  };
  reset();
  for (int j = 0; j < (int)ncol.size(); ++j)
    for (int c0 = 0; c0 < ncol[j]; c0 += 4) {
      const bool newq = cur.nq == 0 || cur.jq[cur.nq - 1] != j;
      if (nslot == 4 || (newq && cur.nq == qmax)) flush();
      if (cur.nq == 0 || cur.jq[cur.nq - 1] != j) cur.jq[cur.nq++] = j;
      cur.c0[nslot] = c0;
      cur.rt[nslot] = cur.nq - 1;
      ++nslot;
    }
  flush();
  return nrt;
}

void eri_run(Context &C, const EriRequest &R) {
  init_md_table();
  std::vector<PairDesc> rd = *R.rdesc, cd = *R.cdesc;
  // ---- pair tables: look up the basis-level cache (a dense shell-pair
  // index, or a hash map for very large bases), build the missing ones
  std::vector<EtJob> jobs;
  const size_t nhdr0 = C.md_hdr_host.size();
  const bool dense = !C.md_pt_dense.empty();
  auto pair_entry = [&](PairDesc &d) {
    if (d.s0 < 0 || d.s0 >= C.nshells || d.s1 < 0 || d.s1 >= C.nshells) throw Error{8, "shell index out of range"};
    const long long key = (long long)d.s0 * C.nshells + d.s1;
    if (dense) {
      const int pt = C.md_pt_dense[(size_t)key];
      if (pt >= 0) {
        d.pt = pt;
        return;
      }
    } else {
      auto it = C.md_pt.find(key);
      if (it != C.md_pt.end()) {
        d.pt = it->second;
        return;
      }
    }
    const int l0 = C.shell_l[d.s0], l1 = C.shell_l[d.s1];
    if (l0 + l1 > 2 * MAXL) {
      char buf[160];
      snprintf(buf, sizeof buf, "no ERI kernel for a shell pair with l0 + l1 = %d (> %d)", l0 + l1, 2 * MAXL);
      throw Error{3, buf};
    }
    const int n0 = C.shell_nang[d.s0], n1 = C.shell_nang[d.s1];
    const int np0 = C.shell_np[d.s0], np1 = C.shell_np[d.s1];
    PairHdr h{n0, n1, np0 * np1, (int)C.md_prim_used};
    for (int ib = 0; ib < np1; ++ib)
      for (int ia = 0; ia < np0; ++ia) {
        jobs.push_back(EtJob{d.s0, d.s1, ia, ib, (int)C.md_prim_used++, 0, C.md_et_used});
        C.md_et_used += (long long)nherm(l0 + l1) * n0 * n1;
      }
    d.pt = (int)C.md_hdr_host.size();
    if (dense)
      C.md_pt_dense[(size_t)key] = d.pt;
    else
      C.md_pt[key] = d.pt;
    C.md_hdr_host.push_back(h);
  };
  for (int r : *R.rows) pair_entry(rd[r]);
  for (int c : *R.cols) pair_entry(cd[c]);
  // rows (orbital pairs), primitive pairs and E = l0 + l1 of every pair used
  std::vector<int> rnr(rd.size(), 0), rnp(rd.size(), 0), rE(rd.size(), 0);
  std::vector<int> cnr(cd.size(), 0), cnp(cd.size(), 0), cE(cd.size(), 0);
  auto describe = [&](const std::vector<PairDesc> &d, const std::vector<int> &use, std::vector<int> &nr,
                      std::vector<int> &np, std::vector<int> &E, const char *what) {
    for (int k : use) {
      const PairDesc &x = d[k];
      nr[k] = pair_rows_host(x, C.shell_nang[x.s0], C.shell_nang[x.s1]);
      if (nr[k] <= 0) throw Error{8, std::string("empty ") + what + " pair"};
      np[k] = C.shell_np[x.s0] * C.shell_np[x.s1];
      E[k] = C.shell_l[x.s0] + C.shell_l[x.s1];
    }
  };
  describe(rd, *R.rows, rnr, rnp, rE, "bra");
  describe(cd, *R.cols, cnr, cnp, cE, "ket");
  if (R.rows->empty() || R.cols->empty()) return;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (C.time_eri) {
    GI_HIP(hipEventCreate(&e0));
    GI_HIP(hipEventCreate(&e1));
  }
  long long nlaunch = 0;
  // device tables that outlive the request (pair headers, primitive records,
  // Hermite tables) grow first: the launch arguments below point into them
  if (C.md_hdr_host.size() != nhdr0) {
    C.md_hdr.alloc(C.md_hdr_host.size());
    C.h2d(C.md_hdr.p, C.md_hdr_host.data(), C.md_hdr_host.size() * sizeof(PairHdr));
  }
  if (!jobs.empty()) {
    C.md_prim.grow_keep((size_t)C.md_prim_used, C.stream);
    C.md_et.grow_keep((size_t)C.md_et_used, C.stream);
  }

  // ---- segments: lists of pairs packed into consecutive records.  Products
  // bin rows by E and columns by F, one segment per non-empty bin, shared by
  // every (E, F) launch of the bin; zipped lists give each (E, F) class its
  // own row and column segments (pairs matched by position).
  constexpr int NB = 2 * MAXL + 1;
  struct Seg {
    int side;               // 0: rows (rd), 1: columns (cd)
    std::vector<int> idx;   // descriptor indices
    size_t off = 0;         // first record
    // aggregates over the segment: primitive pairs, rows, rows x primitive
    // pairs, primitive pairs x ceil(rows / 4); every pair uncontracted / one row
    double s_np = 0, s_nr = 0, s_nrnp = 0, s_r4np = 0;
    bool plain = true, nr1 = true;
  };
  std::vector<Seg> segs;
  struct Launch {
    int E, F;
    int rs, cs;   // row and column segments
  };
  std::vector<Launch> launches;
  if (R.mode == LAUNCH_PRODUCT) {
    int rbin[NB], cbin[NB];
    for (int e = 0; e < NB; ++e) rbin[e] = cbin[e] = -1;
    for (int r : *R.rows) {
      int &b = rbin[rE[r]];
      if (b < 0) {
        b = (int)segs.size();
        segs.push_back(Seg{0, {}});
      }
      segs[b].idx.push_back(r);
    }
    for (int c : *R.cols) {
      int &b = cbin[cE[c]];
      if (b < 0) {
        b = (int)segs.size();
        segs.push_back(Seg{1, {}});
      }
      segs[b].idx.push_back(c);
    }
    for (int e = 0; e < NB; ++e)
      for (int f = 0; f < NB; ++f)
        if (rbin[e] >= 0 && cbin[f] >= 0) launches.push_back(Launch{e, f, rbin[e], cbin[f]});
  } else {
    int zb[NB][NB];
    for (int e = 0; e < NB; ++e)
      for (int f = 0; f < NB; ++f) zb[e][f] = -1;
    for (size_t i = 0; i < R.rows->size(); ++i) {
      const int r = (*R.rows)[i], c = (*R.cols)[i];
      int &z = zb[rE[r]][cE[c]];
      if (z < 0) {
        z = (int)launches.size();
        launches.push_back(Launch{rE[r], cE[c], (int)segs.size(), (int)segs.size() + 1});
        segs.push_back(Seg{0, {}});
        segs.push_back(Seg{1, {}});
      }
      segs[launches[z].rs].idx.push_back(r);
      segs[launches[z].cs].idx.push_back(c);
    }
  }
  std::vector<int> idx;
  std::vector<unsigned char> side;
  for (Seg &g : segs) {
    const std::vector<int> &nr = g.side ? cnr : rnr, &np = g.side ? cnp : rnp;
    for (int k : g.idx) {
      const double n = nr[k], p = np[k];
      g.s_np += p;
      g.s_nr += n;
      g.s_nrnp += n * p;
      g.s_r4np += p * (double)((nr[k] + 3) / 4);
      g.plain = g.plain && np[k] == 1;
      g.nr1 = g.nr1 && nr[k] == 1;
    }
    g.off = idx.size();
    idx.insert(idx.end(), g.idx.begin(), g.idx.end());
    side.insert(side.end(), g.idx.size(), (unsigned char)g.side);
  }
  C.md_rec.alloc(idx.size());

  static const bool stats = getenv("GIGPU_ERI_STATS") != nullptr;
  long long nq_total = 0;
  double flops_total = 0;
  // prepared class launches, dispatched after all are known (largest first,
  // each on the stream with the least work so far)
  struct Prepared {
    MdArgs a;
    long long grid;
    int kind;   // 0 MFMA kernel, 1 lane kernel, 2 runtime-order kernel, 3 4x4x4 kernel, 4 its thin-ket form
    int KE, KF, E, F;
    long long nq;
    double fl;
    long long ct0, nct;   // kind 3: column tiles (offset into ctiles)
    bool plain;           // kind 3: every pair of the launch uncontracted
  };
  std::vector<Prepared> prep;
  std::vector<ColTile> ctiles;
  // column tiles of a segment as the 4x4x4 kernel's ket side, per qmax
  struct Tiles {
    std::vector<ColTile> t;
    long long nrt = 0;
  };
  std::map<std::pair<int, int>, Tiles> tile_cache;
  auto seg_tiles = [&](int sg, int qmax) -> const Tiles & {
    auto key = std::make_pair(sg, qmax);
    auto it = tile_cache.find(key);
    if (it != tile_cache.end()) return it->second;
    const Seg &g = segs[sg];
    const std::vector<int> &nr = g.side ? cnr : rnr;
    std::vector<int> nc;
    nc.reserve(g.idx.size());
    for (int k : g.idx) nc.push_back(nr[k]);
    Tiles T;
    T.nrt = build_col_tiles(nc, qmax, T.t);
    return tile_cache.emplace(key, std::move(T)).first->second;
  };
  for (size_t i = 0; i < launches.size(); ++i) {
    const Launch &L = launches[i];
    const Seg &SR = segs[L.rs], &SC = segs[L.cs];
    long long nrows = (long long)SR.idx.size(), ncols = (long long)SC.idx.size();
    if (nrows == 0 || ncols == 0) continue;
    const long long nq = (R.mode == LAUNCH_PRODUCT) ? nrows * ncols : nrows;
    const bool plain = SR.plain && SC.plain;
    // lane-per-quartet kernel for small uncontracted classes
    // the lane kernels want the thin side (lower Hermite order) as their bra:
    // a product with E > F runs transposed on the (F, E) kernel
    bool trans = false;
    bool lane_path = false;
    // classes beyond the templated kernels (a pair with l0 + l1 > 6) take the
    // runtime-order kernel, one quartet per workgroup, never transposed
    const bool generic = std::max(L.E, L.F) > MD_EMAX;
    if (!generic && C.eri_lane && std::max(L.E, L.F) <= C.eri_lane_emax) {
      const bool lt = R.mode == LAUNCH_PRODUCT && C.eri_trans && L.E > L.F && lane_table[L.F][L.E] != nullptr;
      lane_path = lt || lane_table[L.E][L.F] != nullptr;
      trans = lt;
    }
    // measured per class on the (H2O)16 decomposition (caller's E rows, F
    // columns): the lane kernels win where both sides are thin; the 4x4x4
    // kernel wins for (2,1), (2,2), (2,3), (3,1), (3,2), (0,4), (4,0)
    if (C.eri_md4 && lane_path) {
      const int e = L.E, f = L.F;
      const bool lane_wins = (e <= 1 && f <= 3) || (f == 0 && e <= 3) || (e == 2 && f == 0);
      if (!lane_wins) lane_path = false, trans = false;
    }
    lane_path = lane_path && plain;
    if (!lane_path) trans = false;
    // s-s side: the thin-ket 4x4x4 kernel with the s-s pairs as its ket (one
    // column each, sixteen quartets per instruction)
    bool thin = false;
    if (!generic && R.mode == LAUNCH_PRODUCT && C.eri_md4 && std::min(L.E, L.F) == 0 &&
        std::max(L.E, L.F) >= C.eri_thin_emin) {
      const bool tt = L.F != 0;   // E = 0 < F: the caller's rows are the s-s side
      thin = (!tt || C.eri_trans) && (tt ? SR.nr1 : SC.nr1);
      if (thin) {
        lane_path = false;
        trans = tt;
      }
    }
    // MFMA kernel orientation: its bra side (stage-2 rows, 16-row tiles) is
    // the side with more rows per pair; the thin side is stacked in columns
    if (!thin && !lane_path && !generic && R.mode == LAUNCH_PRODUCT && C.eri_trans)
      trans = SC.s_nr / (double)ncols > SR.s_nr / (double)nrows;
    // 4x4x4 kernel (products): orientation and column tiles from its padded
    // instruction count -- per tile and bra primitive pair, NHG (KSF ket
    // primitives + ceil(nr / 4)) instructions, plus the R tables (about one
    // instruction's time per 40 recursion entries)
    const bool use4 = !thin && !lane_path && !generic && R.mode == LAUNCH_PRODUCT && C.eri_md4;
    const Tiles *t4 = nullptr;
    if (use4) {
      double best = 0;
      for (int o = 0; o < 2; ++o) {
        if (o == 1 && !C.eri_trans) break;
        const int e = o ? L.F : L.E, f = o ? L.E : L.F;
        const Seg &bra = o ? SC : SR, &ket = o ? SR : SC;
        const Tiles &T = seg_tiles(o ? L.rs : L.cs, m4_qmax(e + f));
        const double np_c = ket.s_np / (double)std::max<size_t>(1, ket.idx.size());
        const double nhg = (nherm(e) + 3) / 4, ksf = (nherm(f) + 3) / 4;
        double rentries = 0;
        for (int m = 0; m <= e + f; ++m) rentries += nherm(m);
        const double cost = (double)T.t.size() * nhg * (ksf * np_c * bra.s_np + bra.s_r4np) +
                            (double)T.nrt * np_c * bra.s_np * rentries / 40.0;
        if (o == 0 || cost < best) {
          best = cost;
          trans = o == 1;
          t4 = &T;
        }
      }
    }
    const int KE = trans ? L.F : L.E, KF = trans ? L.E : L.F;   // kernel bra / ket orders
    const Seg &KR = trans ? SC : SR, &KC = trans ? SR : SC;
    if (trans) std::swap(nrows, ncols);
    // algorithmic operation count in the kernel's own roles (useful rows/columns only)
    const double NHE = nherm(KE), NHF = nherm(KF), rt = rtable_flops(L.E + L.F);
    double fl = 0;
    if (R.mode == LAUNCH_PRODUCT) {
      if (lane_path)   // 2 nr NHF (NHE + nc) per quartet
        fl = 2.0 * NHF * (NHE * KR.s_nr * (double)ncols + KR.s_nr * KC.s_nr) + rt * (double)nrows * (double)ncols;
      else             // stage 1 2 NHE NHF nc per primitive quartet, stage 2 2 nr NHE nc per bra primitive pair
        fl = 2.0 * NHE * NHF * KR.s_np * KC.s_nrnp + 2.0 * NHE * KR.s_nrnp * KC.s_nr + rt * KR.s_np * KC.s_np;
    } else {
      const std::vector<int> &kr = KR.idx, &kc = KC.idx;
      for (size_t k = 0; k < kr.size(); ++k) {
        const double nr = rnr[kr[k]], pr = rnp[kr[k]];
        const double nc = cnr[kc[k]], pc = cnp[kc[k]];
        if (lane_path)
          fl += 2.0 * nr * NHF * (NHE + (R.mode == LAUNCH_DIAG ? 1.0 : nc)) + rt;
        else
          fl += 2.0 * NHE * NHF * nc * pr * pc + (R.mode == LAUNCH_DIAG ? 2.0 * NHE * nc * pr : 2.0 * NHE * nr * nc * pr) +
                rt * pr * pc;
      }
    }
    MdArgs a{};
    a.prim = C.md_prim.p;
    a.et = C.md_et.p;
    a.rrec = C.md_rec.p + KR.off;
    a.crec = C.md_rec.p + KC.off;
    a.nrows = nrows;
    a.ncols = ncols;
    a.mode = R.mode;
    a.trans = trans ? 1 : 0;
    a.kt = 1;
    a.kappa = R.kappa;
    a.out = R.out;
    a.ldo = R.ldo;
    a.sel = R.d_sel;
    a.boys = C.dboys.p;
    long long grid;
    long long ct0 = 0, nct = 0;
    // 4x4x4 kernels: row blocks of M4_NB bra pairs, or fewer when that leaves
    // fewer than eri_md4_tiles tiles (small launches: more, shorter wavefronts)
    auto row_block = [&](long long ncolt) {
      int nb = M4_NB;
      while (nb > 1 && (nrows + nb - 1) / nb * ncolt < C.eri_md4_tiles) nb /= 2;
      return nb;
    };
    if (thin) {
      const long long ncolt = (ncols + M4S_NK - 1) / M4S_NK;
      a.kt = row_block(ncolt);
      a.ntiles = (nrows + a.kt - 1) / a.kt * ncolt;
      grid = std::min<long long>(a.ntiles, 8192);
      if (grid >= 8) grid = (grid + 7) / 8 * 8;
    } else if (use4) {
      ct0 = (long long)ctiles.size();
      nct = (long long)t4->t.size();
      ctiles.insert(ctiles.end(), t4->t.begin(), t4->t.end());
      a.kt = row_block(nct);
      a.ntiles = (nrows + a.kt - 1) / a.kt * nct;   // row blocks x column tiles
      grid = std::min<long long>(a.ntiles, 8192);
      if (grid >= 8) grid = (grid + 7) / 8 * 8;
    } else if (lane_path) {
      const long long ng = (nrows + 63) / 64;
      const long long nt = (R.mode == LAUNCH_PRODUCT) ? ng * ncols : ng;
      a.ntiles = nt;
      grid = std::min<long long>((nt + 3) / 4, 16384);
    } else {
      if (R.mode == LAUNCH_PRODUCT) {
        // enough ket pairs per tile to fill 32 stacked columns, while keeping
        // at least eri_min_tiles tiles when the launch is large enough
        long long kt = (long long)std::ceil(32.0 / std::max(1.0, KC.s_nr / (double)ncols));
        kt = std::max<long long>(1, std::min<long long>(kt, C.eri_max_kt));
        while (kt > 1 && nrows * ((ncols + kt - 1) / kt) < C.eri_min_tiles) kt = (kt + 1) / 2;
        a.kt = (int)kt;
        a.ntiles = nrows * ((ncols + a.kt - 1) / a.kt);
      } else {
        a.ntiles = nrows;
      }
      grid = std::min<long long>(a.ntiles, 4096);
      if (grid >= 8) grid = (grid + 7) / 8 * 8;
    }
    prep.push_back(Prepared{a, grid, thin ? 4 : (use4 ? 3 : (generic ? 2 : (lane_path ? 1 : 0))), KE, KF, L.E, L.F,
                            nq, fl, ct0, nct, plain});
    nq_total += nq;
    flops_total += fl;
  }
  // the request's host tables (pair-table jobs, descriptors, launch index
  // lists, column tiles) go to the device in one copy from pinned memory;
  // the kernels read them from that one device block
  const EtJob *d_jobs;
  const PairDesc *d_rd, *d_cd;
  const int *d_idx;
  const unsigned char *d_side;
  const ColTile *d_ct;
  {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_jobs = 0;
    const size_t o_rd = o_jobs + al(jobs.size() * sizeof(EtJob));
    const size_t o_cd = o_rd + al(rd.size() * sizeof(PairDesc));
    const size_t o_idx = o_cd + al(cd.size() * sizeof(PairDesc));
    const size_t o_side = o_idx + al(idx.size() * sizeof(int));
    const size_t o_ct = o_side + al(side.size());
    const size_t nbytes = o_ct + al(ctiles.size() * sizeof(ColTile));
    unsigned char *h = C.stage_host(nbytes);
    auto put = [&](size_t o, const void *src, size_t n) {
      if (n) memcpy(h + o, src, n);
    };
    put(o_jobs, jobs.data(), jobs.size() * sizeof(EtJob));
    put(o_rd, rd.data(), rd.size() * sizeof(PairDesc));
    put(o_cd, cd.data(), cd.size() * sizeof(PairDesc));
    put(o_idx, idx.data(), idx.size() * sizeof(int));
    put(o_side, side.data(), side.size());
    put(o_ct, ctiles.data(), ctiles.size() * sizeof(ColTile));
    C.md_blob.alloc(nbytes);
    GI_HIP(hipMemcpyAsync(C.md_blob.p, h, nbytes, hipMemcpyHostToDevice, C.stream));
    GI_HIP(hipEventRecord(C.md_stage_ev, C.stream));   // the next stage_host waits for it
    unsigned char *b = C.md_blob.p;
    d_jobs = (const EtJob *)(b + o_jobs);
    d_rd = (const PairDesc *)(b + o_rd);
    d_cd = (const PairDesc *)(b + o_cd);
    d_idx = (const int *)(b + o_idx);
    d_side = (const unsigned char *)(b + o_side);
    d_ct = (const ColTile *)(b + o_ct);
  }
  // the timed window (when asked for) spans the request's kernels only:
  // pair tables, record packing and the class launches
  if (C.time_eri) GI_HIP(hipEventRecord(e0, C.stream));
  if (!jobs.empty()) {
    hipLaunchKernelGGL(k_build_et, dim3((unsigned)jobs.size()), dim3(64), 0, C.stream, C.dshells.p, C.dexps.p,
                       C.dcntr.p, C.dnorms.p, C.spher, d_jobs, C.md_prim.p, C.md_et.p);
    GI_HIP(hipGetLastError());
    ++nlaunch;
  }
  if (!idx.empty()) {
    hipLaunchKernelGGL(k_pack_recs, dim3((unsigned)((idx.size() + 255) / 256)), dim3(256), 0, C.stream, d_rd, d_cd,
                       d_idx, d_side, (long long)idx.size(), C.md_hdr.p, C.md_prim.p, C.md_rec.p);
    GI_HIP(hipGetLastError());   // record packing: not counted among the ERI kernel launches
  }
  // per-class statistics time every launch on its own, so they stay on the
  // context stream; otherwise the launches are spread over the side streams
  const int nstream = stats ? 1 : (int)std::min<long long>(C.eri_streams, (long long)prep.size());
  if (nstream > 1) {
    if (!C.eri_fork) {
      GI_HIP(hipEventCreateWithFlags(&C.eri_fork, hipEventDisableTiming));
      for (int j = 0; j < Context::ERI_MAX_STREAMS - 1; ++j) {
        GI_HIP(hipStreamCreateWithFlags(&C.eri_side[j], hipStreamNonBlocking));
        GI_HIP(hipEventCreateWithFlags(&C.eri_join[j], hipEventDisableTiming));
      }
    }
    GI_HIP(hipEventRecord(C.eri_fork, C.stream));
    for (int j = 1; j < nstream; ++j) GI_HIP(hipStreamWaitEvent(C.eri_side[j - 1], C.eri_fork, 0));
  }
  std::vector<size_t> order(prep.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return prep[x].fl > prep[y].fl; });
  std::vector<double> load(nstream, 0.0);
  for (size_t oi = 0; oi < order.size(); ++oi) {
    const Prepared &P = prep[order[oi]];
    const int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[j] += P.fl;
    hipStream_t st = j == 0 ? C.stream : C.eri_side[j - 1];
    hipEvent_t s0 = nullptr, s1 = nullptr;
    if (stats) {
      GI_HIP(hipEventCreate(&s0));
      GI_HIP(hipEventCreate(&s1));
      GI_HIP(hipEventRecord(s0, st));
    }
    if (P.kind == 4)
      md4s_table[P.KE](P.a, P.plain, (int)P.grid, st);
    else if (P.kind == 3)
      md4_table[P.KE][P.KF](P.a, d_ct + P.ct0, P.nct, P.plain, (int)P.grid, st);
    else if (P.kind == 2)
      launch_generic(P.KE, P.KF, P.a, st);
    else
      (P.kind == 1 ? lane_table[P.KE][P.KF] : md_table[P.KE][P.KF])(P.a, (int)P.grid, st);
    GI_HIP(hipGetLastError());
    ++nlaunch;
    if (stats) {
      GI_HIP(hipEventRecord(s1, st));
      GI_HIP(hipEventSynchronize(s1));
      float ms = 0;
      GI_HIP(hipEventElapsedTime(&ms, s0, s1));
      auto &sc = C.class_stats[P.E * 16 + P.F];
      sc[0] += (double)P.nq;
      sc[1] += ms;
      sc[2] += P.fl;
      (void)hipEventDestroy(s0);
      (void)hipEventDestroy(s1);
    }
  }
  for (int j = 1; j < nstream; ++j) {
    GI_HIP(hipEventRecord(C.eri_join[j - 1], C.eri_side[j - 1]));
    GI_HIP(hipStreamWaitEvent(C.stream, C.eri_join[j - 1], 0));
  }
  if (C.time_eri) {
    GI_HIP(hipEventRecord(e1, C.stream));
    GI_HIP(hipEventSynchronize(e1));
    float ms = 0;
    GI_HIP(hipEventElapsedTime(&ms, e0, e1));
    C.eri_ms_acc += ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  C.eri_quartets_acc += nq_total;
  C.eri_flops_acc += (long long)flops_total;
  C.eri_launches_acc += nlaunch;
  // the pair table and index buffers are reused by the next request
  C.wait();
}

}  // namespace gi
