// Coulomb energy of an AO density and its nuclear gradient on the device:
//   E_J = 1/2 sum_{pqrs} D_pq (pq|rs) D_rs,   dE_J/dR_A for every atom A.
// The reference has no gradient code (its ERIs only feed the Cholesky
// decomposition); this is the differentiation form, built on the same
// Hermite-matrix kernels as the ERIs (md_kernel.h).
//
// With the density folded into each primitive pair's Hermite table,
//   D^P_h = sum_{ab in P} w D_ab ET(h, ab)      (w = 2 for two different shells),
// the energy is a sum over ordered pairs of primitive pairs,
//   E_J = 1/2 sum_{P,Q} f(P, Q),   f = sum_{h,k} D^P_h (-1)^{|k|} D^Q_k R_{h+k}(P - Q),
// and, f being symmetric in P and Q, dE_J/dx = sum_{P,Q} d_P f(P, Q): only the
// bra pair is differentiated.  A pair with centres A, B and exponents a, b
// depends on the coordinates through X_AB = A - B (the E coefficients and
// exp(-ab/p AB^2)) and through P = (aA + bB)/p (R, with dR_tuv/dP_x =
// R_{t+1,u,v}):
//   d_A f = df/dX_AB + (a/p) df/dP,    d_B f = -df/dX_AB + (b/p) df/dP.
// Every primitive pair gets a table of GRAD_NCOL columns over the Hermite
// indices of order <= E + 1 (E = la + lb):
//   column 0     D^P_h                        (zero at order E + 1)
//   columns 1-3  dD^P_h / dX_AB,d, d = x, y, z (differentiated E recursion)
//   columns 4-6  D^P_{h - e_d}                (df/dP_d written as a bra of order E + 1)
//   columns 7-9  (a/p) D^P_{h - e_d}
// One md_kernel launch per (E + 1, F) class pair takes these tables as bra
// rows and column 0 of the other pairs' tables as ket columns, and sums every
// row over all ket pairs (rowsum mode).  No derivative integral is stored: the
// R tables of order E + F + 1 are recomputed per primitive quartet.
#include <algorithm>
#include <cmath>
#include "gi_context.h"

namespace gi {

// launch records of the shell pairs: first primitive, table, destination rows
__global__ void k_grad_recs(const PrimRec *__restrict__ prims, const int *__restrict__ prim0,
                            const int *__restrict__ nprim, long long n, int nrow, QRec *out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PrimRec pr = prims[prim0[i]];
  out[i] = QRec{pr.px, pr.py, pr.pz, pr.p, pr.et, GRAD_NCOL * i, nrow, 0, 0, GRAD_NCOL, 1, nprim[i], prim0[i], 0};
}

// shell pairs (s0 >= s1, s1-major order) kept by the screening of
// coulomb_energy_grad: the Schwarz diagonals (ab|ab) of every pair from one
// diagonal ERI request, the density maxima on the host
static std::vector<char> screened_pairs(Context &C, const double *D, long long ldd, double kappa, double eps) {
  const int nsh = C.nshells;
  std::vector<PairDesc> rd, cd;
  std::vector<int> rows;
  long long loc = 0;
  for (int s1 = 0; s1 < nsh; ++s1)
    for (int s0 = s1; s0 < nsh; ++s0) {
      const int fl = s0 == s1 ? PAIR_SAME : 0;
      rd.push_back(PairDesc{s0, s1, fl, 0, 0, 0, loc});
      cd.push_back(PairDesc{s0, s1, fl, 0, 0, 0, 0});
      rows.push_back((int)rows.size());
      const int n0 = C.shell_nang[s0], n1 = C.shell_nang[s1];
      loc += s0 == s1 ? (long long)n0 * (n0 + 1) / 2 : (long long)n0 * n1;
    }
  DBuf<double> dq;
  dq.alloc((size_t)loc);
  EriRequest R;
  R.rdesc = &rd;
  R.cdesc = &cd;
  R.rows = &rows;
  R.cols = &rows;
  R.mode = LAUNCH_DIAG;
  R.out = dq.p;
  R.ldo = 0;
  R.kappa = kappa;
  eri_run(C, R);
  std::vector<double> q((size_t)loc);
  C.read(q.data(), dq.p, sizeof(double) * (size_t)loc);
  std::vector<double> w(rows.size());
  double wsum = 0.0;
  for (size_t i = 0; i < rows.size(); ++i) {
    const int s0 = rd[i].s0, s1 = rd[i].s1;
    const int n0 = C.shell_nang[s0], n1 = C.shell_nang[s1];
    const long long o = rd[i].off, nr = s0 == s1 ? (long long)n0 * (n0 + 1) / 2 : (long long)n0 * n1;
    double qm = 0.0, dm = 0.0;
    for (long long r = 0; r < nr; ++r) qm = std::max(qm, q[(size_t)(o + r)]);
    for (int b = 0; b < n1; ++b)
      for (int a = 0; a < n0; ++a) {
        const long long pa = C.shell_loc[s0] + a, pb = C.shell_loc[s1] + b;
        dm = std::max(dm, std::max(std::fabs(D[pa + ldd * pb]), std::fabs(D[pb + ldd * pa])));
      }
    w[i] = std::sqrt(std::max(qm, 0.0)) * dm;
    wsum += w[i];
  }
  std::vector<char> keep(rows.size());
  for (size_t i = 0; i < rows.size(); ++i) keep[i] = w[i] * wsum >= eps ? 1 : 0;
  return keep;
}

void coulomb_energy_grad(Context &C, const double *D, long long ldd, double kappa, double *energy,
                         double *grad) {
  if (!C.basis_set) throw Error{4, "basis not set"};
  if (!D || !energy) throw Error{61, "null density or energy pointer"};
  if (ldd < C.nao) throw Error{61, "leading dimension of D smaller than the number of AOs"};
  const bool want_grad = grad != nullptr;
  const int nsh = C.nshells;
  // shell pairs s0 >= s1, grouped by E = l0 + l1 so that a class is a contiguous range
  const int nbin = 2 * MAXL + 1;
  std::vector<std::vector<std::pair<int, int>>> bins(nbin);
  // R tables of order (E + 1) + F for the gradient: within the Boys table
  if (want_grad && 4 * C.lmax + 1 > BOYS_MAXM)
    throw Error{62, "the nuclear gradient needs 4 lmax + 1 <= " + std::to_string(BOYS_MAXM) + " (l <= 4)"};
  // density-weighted Schwarz screening (grad_screen = v > 0, eps = 10^-v): with
  // w_s = max_{ab in s} |D_ab| sqrt(max_{ab in s} (ab|ab)) for shell pair s,
  // every term of E_J that involves s is at most w_s sum_t w_t, and the pair
  // is left out when that is below eps (off by default: every pair kept)
  std::vector<char> keep;
  if (C.grad_screen > 0) keep = screened_pairs(C, D, ldd, kappa, std::pow(10.0, -(double)C.grad_screen));
  long long ip = 0;
  for (int s1 = 0; s1 < nsh; ++s1)
    for (int s0 = s1; s0 < nsh; ++s0, ++ip)
      if (keep.empty() || keep[ip]) bins[C.shell_l[s0] + C.shell_l[s1]].push_back({s0, s1});
  std::vector<long long> bin_off(nbin + 1, 0);
  for (int e = 0; e < nbin; ++e) bin_off[e + 1] = bin_off[e] + (long long)bins[e].size();
  const long long npair = bin_off[nbin];
  std::vector<GradJob> jobs;
  std::vector<int> prim0(npair), nprim(npair), pa(npair), pb(npair);
  long long etn = 0;
  for (int e = 0; e < nbin; ++e)
    for (size_t k = 0; k < bins[e].size(); ++k) {
      const long long i = bin_off[e] + (long long)k;
      const int s0 = bins[e][k].first, s1 = bins[e][k].second;
      pa[i] = s0;
      pb[i] = s1;
      prim0[i] = (int)jobs.size();
      nprim[i] = C.shell_np[s0] * C.shell_np[s1];
      for (int ib = 0; ib < C.shell_np[s1]; ++ib)
        for (int ia = 0; ia < C.shell_np[s0]; ++ia) {
          GradJob j{};
          j.s0 = s0;
          j.s1 = s1;
          j.ia = ia;
          j.ib = ib;
          j.prim = (int)jobs.size();
          j.et = etn;
          j.w = s0 != s1 ? 2.0 : 1.0;
          jobs.push_back(j);
          etn += (long long)nherm(e + 1) * GRAD_NCOL;
        }
    }
  const long long nao = C.nao;
  std::vector<int> aoloc(C.shell_loc.begin(), C.shell_loc.end());
  DBuf<double> dD, det, drows;
  DBuf<GradJob> djobs;
  DBuf<PrimRec> dprim;
  DBuf<int> daoloc, dprim0, dnprim;
  DBuf<QRec> dbra, dket;
  dD.upload(D, (size_t)ldd * (size_t)(nao - 1) + (size_t)nao, C.stream);
  daoloc.upload(aoloc, C.stream);
  djobs.upload(jobs, C.stream);
  dprim0.upload(prim0, C.stream);
  dnprim.upload(nprim, C.stream);
  det.alloc((size_t)etn);
  dprim.alloc(jobs.size());
  launch_build_grad_et(C, daoloc.p, dD.p, ldd, djobs.p, (long long)jobs.size(), dprim.p, det.p);
  dbra.alloc(npair);
  dket.alloc(npair);
  const unsigned gb = (unsigned)((npair + 255) / 256);
  hipLaunchKernelGGL(k_grad_recs, dim3(gb), dim3(256), 0, C.stream, dprim.p, dprim0.p, dnprim.p, npair,
                     want_grad ? GRAD_NCOL : 1, dbra.p);
  hipLaunchKernelGGL(k_grad_recs, dim3(gb), dim3(256), 0, C.stream, dprim.p, dprim0.p, dnprim.p, npair, 1, dket.p);
  drows.alloc((size_t)npair * GRAD_NCOL);
  GI_HIP(hipMemsetAsync(drows.p, 0, sizeof(double) * (size_t)npair * GRAD_NCOL, C.stream));
  for (int eb = 0; eb < nbin; ++eb) {
    const long long nr = bin_off[eb + 1] - bin_off[eb];
    if (nr == 0) continue;
    for (int fk = 0; fk < nbin; ++fk) {
      const long long nc = bin_off[fk + 1] - bin_off[fk];
      if (nc == 0) continue;
      MdArgs a{};
      a.prim = dprim.p;
      a.et = det.p;
      a.rrec = dbra.p + bin_off[eb];
      a.crec = dket.p + bin_off[fk];
      a.nrows = nr;
      a.ncols = nc;
      a.mode = LAUNCH_PRODUCT;
      a.trans = 0;
      a.rowsum = 1;
      a.kappa = kappa;
      a.out = drows.p;
      a.ldo = 0;
      a.sel = nullptr;
      a.boys = C.dboys.p;
      // one-column ket pairs: 32 of them fill the stacked stage-2 columns
      long long kt = std::max<long long>(1, std::min<long long>(32, C.eri_max_kt));
      while (kt > 1 && nr * ((nc + kt - 1) / kt) < C.eri_min_tiles) kt = (kt + 1) / 2;
      a.kt = (int)kt;
      a.ntiles = nr * ((nc + kt - 1) / kt);
      long long grid = std::min<long long>(a.ntiles, 4096);
      if (grid >= 8) grid = (grid + 7) / 8 * 8;
      launch_md_class(want_grad ? eb + 1 : eb, fk, a, (int)grid, C.stream);
      GI_HIP(hipGetLastError());
    }
  }
  std::vector<double> rows((size_t)npair * GRAD_NCOL);
  GI_HIP(hipMemcpyAsync(rows.data(), drows.p, sizeof(double) * rows.size(), hipMemcpyDeviceToHost, C.stream));
  GI_HIP(hipStreamSynchronize(C.stream));
  double e = 0.0;
  for (long long i = 0; i < npair; ++i) e += rows[GRAD_NCOL * i];
  *energy = 0.5 * e;
  if (want_grad) {
    std::fill(grad, grad + 3 * (size_t)C.natoms, 0.0);
    for (long long i = 0; i < npair; ++i) {
      const int A = C.shell_atom[pa[i]], B = C.shell_atom[pb[i]];
      const double *r = rows.data() + GRAD_NCOL * i;
      for (int d = 0; d < 3; ++d) {
        const double gE = r[1 + d], gR = r[4 + d], gRa = r[7 + d];
        grad[3 * A + d] += gE + gRa;
        grad[3 * B + d] += -gE + (gR - gRa);
      }
    }
  }
}

}  // namespace gi
