// C ABI of libgigpu (declarations and reference citations in include/gigpu.h).
#include "../../include/gigpu.h"
#include <atomic>
#include <cstring>
#include <memory>
#include <thread>
#include "gi_context.h"

namespace gi {
void chol_build(Context &C, double tau, double kappa, long long max_block_dim, long long block_dim,
                long long flags);
void mo_transform(Context &C, const double *dCA, long long lda, long long a0, long long a1, const double *dCB,
                  long long ldb, long long b0, long long b1, long long max_buffer_mb, long long flags,
                  double *dRkab);
}  // namespace gi

struct gi_context {
  std::unique_ptr<gi::Context> c;
};

using gi::Error;

#define GI_TRY(...)                                    \
  try {                                                \
    __VA_ARGS__;                                       \
    return 0;                                          \
  } catch (const gi::Error &e) {                       \
    gi::set_error(e.msg);                              \
    return e.code ? e.code : 1;                        \
  } catch (const std::exception &e) {                  \
    gi::set_error(std::string("exception: ") + e.what()); \
    return 99;                                         \
  }

static gi::Context &ctx_of(gi_context *c) {
  if (!c || !c->c) throw Error{7, "null gi_context"};
  GI_HIP(hipSetDevice(c->c->device));
  gi::g_alloc_stream = c->c->stream;
  return *c->c;
}

extern "C" {

const char *gi_last_error(void) { return gi::last_error(); }

int gi_create(gi_context **out, int64_t device) {
  GI_TRY({
    if (!out) throw Error{7, "null output pointer"};
    auto *c = new gi_context;
    try {
      c->c = std::make_unique<gi::Context>((int)device);
      GI_HIP(hipSetDevice((int)device));
      gi::noalloc_setup((int)device);
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  })
}

int gi_destroy(gi_context *c) {
  GI_TRY({ delete c; })
}

int gi_set_basis(gi_context *cp, int64_t natoms, const double *coords, int64_t nshells,
                 const int64_t *shell_centers, const int64_t *shell_params, int64_t nparams,
                 const int64_t *momentum, const int64_t *nprim, int64_t max_nprim, const double *cntr,
                 const double *exps, int64_t max_nang, const double *norm, int64_t spher) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (natoms <= 0 || nshells <= 0 || nparams <= 0 || max_nprim <= 0 || max_nang <= 0)
      throw Error{8, "invalid basis dimensions"};
    int lmax = 0;
    for (int64_t p = 0; p < nparams; ++p) {
      if (momentum[p] < 0 || momentum[p] > gi::MAXL)
        throw Error{9, "Basis set includes angular momenta unsupported by the Auto2e subroutine"};
      if (nprim[p] < 1 || nprim[p] > max_nprim) throw Error{8, "invalid number of primitives"};
      lmax = std::max<int>(lmax, (int)momentum[p]);
    }
    C.natoms = (int)natoms;
    C.nshells = (int)nshells;
    C.spher = spher ? 1 : 0;
    C.lmax = lmax;
    C.hshells.assign(nshells, gi::DevShell{});
    C.shell_l.assign(nshells, 0);
    C.shell_nang.assign(nshells, 0);
    C.shell_np.assign(nshells, 0);
    C.shell_loc.assign(nshells, 0);
    C.shell_atom.assign(nshells, 0);
    C.atom_xyz.assign(coords, coords + 3 * natoms);
    std::vector<double> hexp, hcn, hnorm;
    int loc = 0;
    for (int64_t s = 0; s < nshells; ++s) {
      int64_t a = shell_centers[s] - 1, p = shell_params[s] - 1;
      if (a < 0 || a >= natoms || p < 0 || p >= nparams) throw Error{8, "shell index out of range"};
      int l = (int)momentum[p];
      int np = (int)nprim[p];
      int nang = C.spher ? (l >= 2 ? 2 * l + 1 : gi::ncart(l)) : gi::ncart(l);
      if (nang > max_nang) throw Error{8, "max_nang smaller than a shell's angular function count"};
      gi::DevShell d{};
      d.x = coords[3 * a];
      d.y = coords[3 * a + 1];
      d.z = coords[3 * a + 2];
      d.l = l;
      d.np = np;
      d.prim0 = (int)hexp.size();
      d.nang = nang;
      d.norm0 = (int)hnorm.size();
      for (int i = 0; i < np; ++i) {
        hexp.push_back(exps[i + max_nprim * p]);
        hcn.push_back(cntr[i + max_nprim * p]);
      }
      for (int i = 0; i < nang; ++i) hnorm.push_back(norm[i + max_nang * p]);
      C.hshells[s] = d;
      C.shell_l[s] = l;
      C.shell_nang[s] = nang;
      C.shell_np[s] = np;
      C.shell_loc[s] = loc;
      C.shell_atom[s] = (int)a;
      loc += nang;
    }
    C.nao = loc;
    C.dshells.upload(C.hshells, C.stream);
    C.dexps.upload(hexp, C.stream);
    C.dcntr.upload(hcn, C.stream);
    C.dnorms.upload(hnorm, C.stream);
    GI_HIP(hipStreamSynchronize(C.stream));
    C.basis_set = true;
    C.chol.valid = false;
    C.reset_pair_cache();
  })
}

int gi_eri_quartets(gi_context *cp, int64_t nq, const int64_t *quartets, double kappa, const int64_t *offsets,
                    double *out, int64_t out_len) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (!C.basis_set) throw Error{4, "basis not set"};
    for (int64_t q = 0; q < nq; ++q) {
      int64_t n = 1;
      for (int k = 0; k < 4; ++k) {
        int64_t s = quartets[4 * q + k];
        if (s < 0 || s >= C.nshells) throw Error{8, "shell index out of range"};
        n *= C.shell_nang[s];
      }
      if (offsets[q] < 0 || offsets[q] + n > out_len) throw Error{8, "output block out of range"};
    }
    gi::DBuf<double> dout;
    dout.alloc(out_len);
    GI_HIP(hipMemsetAsync(dout.p, 0, sizeof(double) * out_len, C.stream));
    // group by the bra pair's angular-function count: one leading dimension per group
    std::map<int, std::vector<int64_t>> groups;
    for (int64_t q = 0; q < nq; ++q) {
      int s0 = (int)quartets[4 * q], s1 = (int)quartets[4 * q + 1];
      groups[C.shell_nang[s0] * C.shell_nang[s1]].push_back(q);
    }
    for (auto &g : groups) {
      std::vector<gi::PairDesc> rd, cd;
      std::vector<int> rows;
      for (int64_t q : g.second) {
        rd.push_back(gi::PairDesc{(int)quartets[4 * q], (int)quartets[4 * q + 1], 0, 0, 0, 0, offsets[q]});
        cd.push_back(gi::PairDesc{(int)quartets[4 * q + 2], (int)quartets[4 * q + 3], 0, 0, 0, 0, 0});
        rows.push_back((int)rows.size());
      }
      gi::EriRequest R;
      R.rdesc = &rd; R.cdesc = &cd;
      R.rows = &rows; R.cols = &rows; R.mode = gi::LAUNCH_ZIP;
      R.out = dout.p; R.ldo = g.first; R.kappa = kappa;
      gi::eri_run(C, R);
    }
    GI_HIP(hipMemcpy(out, dout.p, sizeof(double) * out_len, hipMemcpyDeviceToHost));
  })
}

int gi_eri_pair_block(gi_context *cp, int64_t nrp, const int64_t *row_pairs, int64_t ncp, const int64_t *col_pairs,
                      double kappa, double *out, int64_t ldo) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (!C.basis_set) throw Error{4, "basis not set"};
    if (nrp <= 0 || ncp <= 0) throw Error{8, "empty pair list"};
    std::vector<gi::PairDesc> rd, cd;
    std::vector<int> rows, cols;
    long long nr = 0, nc = 0;
    for (int64_t i = 0; i < nrp; ++i) {
      const int64_t a = row_pairs[2 * i], b = row_pairs[2 * i + 1];
      if (a < 0 || a >= C.nshells || b < 0 || b >= C.nshells) throw Error{8, "shell index out of range"};
      rd.push_back(gi::PairDesc{(int)a, (int)b, 0, 0, 0, 0, nr});
      rows.push_back((int)i);
      nr += (long long)C.shell_nang[a] * C.shell_nang[b];
    }
    if (ldo < nr) throw Error{8, "leading dimension smaller than the number of rows"};
    for (int64_t i = 0; i < ncp; ++i) {
      const int64_t a = col_pairs[2 * i], b = col_pairs[2 * i + 1];
      if (a < 0 || a >= C.nshells || b < 0 || b >= C.nshells) throw Error{8, "shell index out of range"};
      cd.push_back(gi::PairDesc{(int)a, (int)b, 0, 0, 0, 0, ldo * nc});
      cols.push_back((int)i);
      nc += (long long)C.shell_nang[a] * C.shell_nang[b];
    }
    gi::DBuf<double> dout;
    dout.alloc((size_t)ldo * nc);
    GI_HIP(hipMemsetAsync(dout.p, 0, sizeof(double) * ldo * nc, C.stream));
    gi::EriRequest R;
    R.rdesc = &rd; R.cdesc = &cd;
    R.rows = &rows; R.cols = &cols; R.mode = gi::LAUNCH_PRODUCT;
    R.out = dout.p; R.ldo = ldo; R.kappa = kappa;
    gi::eri_run(C, R);
    GI_HIP(hipMemcpy(out, dout.p, sizeof(double) * ldo * nc, hipMemcpyDeviceToHost));
  })
}

int gi_set_option(gi_context *cp, const char *name, int64_t value) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    const std::string n = name ? name : "";
    if (n == "eri_min_tiles") C.eri_min_tiles = std::max<int64_t>(1, value);
    else if (n == "eri_max_kt") C.eri_max_kt = std::max<int64_t>(1, value);
    else if (n == "eri_lane") C.eri_lane = value != 0;
    else if (n == "eri_trans") C.eri_trans = value != 0;
    else if (n == "eri_md4") C.eri_md4 = value != 0;
    else if (n == "eri_thin_emin") C.eri_thin_emin = std::max<int64_t>(0, value);
    else if (n == "eri_md4_tiles") C.eri_md4_tiles = std::max<int64_t>(0, value);
    else if (n == "micro_persist_max") C.micro_persist_max = std::max<int64_t>(0, value);
    else if (n == "grad_screen") C.grad_screen = std::min<int64_t>(30, std::max<int64_t>(0, value));
    else if (n == "thc_schur_mb") C.thc_schur_mb = std::max<int64_t>(0, value);
    else if (n == "thc_cholesky") C.thc_cholesky = value != 0;
    else if (n == "eri_lane_emax") C.eri_lane_emax = std::max<int64_t>(0, value);
    else if (n == "vec_blocks") C.vec_blocks = std::max<int64_t>(1, value);
    else if (n == "eri_streams")
      C.eri_streams = std::min<int64_t>(gi::Context::ERI_MAX_STREAMS, std::max<int64_t>(1, value));
    else throw Error{9, "unknown option " + n};
  })
}

int gi_chol_build_batch(int64_t n, gi_context *const *ctxs, double tau, double kappa, int64_t max_block_dim,
                        int64_t vec_block, int64_t flags, int64_t nworkers, int64_t *status) {
  GI_TRY({
    if (n < 0) throw Error{8, "negative batch size"};
    if (n == 0) return 0;
    if (!ctxs || !status) throw Error{8, "null batch arrays"};
    if (!(tau > 0.0)) throw Error{6, "tau must be positive"};
    if (max_block_dim <= 0 || vec_block <= 0) throw Error{5, "MaxBlockDim and CholVecsBlock must be positive"};
    std::atomic<int64_t> next{0};
    std::vector<std::string> msgs(n);
    auto worker = [&]() {
      for (;;) {
        const int64_t i = next.fetch_add(1);
        if (i >= n) break;
        try {
          gi::Context &C = ctx_of(ctxs[i]);
          // the molecules of the batch already run concurrently (one stream
          // each); forking every context's ERI launches over side streams as
          // well would oversubscribe the hardware queues.  Its waits sleep
          // instead of spinning: every worker thread waits on the GPU at once
          const long long streams = C.eri_streams;
          C.eri_streams = 1;
          C.blocking_waits = true;
          try {
            gi::chol_build(C, tau, kappa, max_block_dim, vec_block, flags);
          } catch (...) {
            C.eri_streams = streams;
            C.blocking_waits = false;
            throw;
          }
          C.eri_streams = streams;
          C.blocking_waits = false;
          status[i] = 0;
        } catch (const gi::Error &e) {
          status[i] = e.code ? e.code : 1;
          msgs[i] = e.msg;
        } catch (const std::exception &e) {
          status[i] = 99;
          msgs[i] = std::string("exception: ") + e.what();
        }
      }
    };
    const int64_t nw = std::max<int64_t>(1, std::min<int64_t>(nworkers, n));
    std::vector<std::thread> pool;
    for (int64_t w = 1; w < nw; ++w) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    for (int64_t i = 0; i < n; ++i)
      if (status[i] != 0) throw Error{(int)status[i], "batch entry " + std::to_string(i) + ": " + msgs[i]};
  })
}

int gi_coulomb_energy_grad(gi_context *cp, const double *density, int64_t ldd, double kappa, double *energy,
                           double *grad) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    gi::coulomb_energy_grad(C, density, ldd, kappa, energy, grad);
  })
}

int gi_one_electron(gi_context *cp, const double *atom_charges, int64_t npc, const double *pc_coords,
                    const double *pc_charges, double *S, double *T, double *V, int64_t ldh) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    gi::one_electron(C, atom_charges, npc, pc_coords, pc_charges, S, T, V, ldh);
  })
}

int gi_h0(gi_context *cp, const double *atom_charges, int64_t npc, const double *pc_coords, const double *pc_charges,
          double *H0, int64_t ldh) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (!H0) throw Error{8, "null H0"};
    const long long n = C.nao;
    if (ldh < n) throw Error{8, "leading dimension smaller than the number of AOs"};
    // T is written with the caller's leading dimension, like H0
    std::vector<double> Tm((size_t)ldh * (n - 1) + n);
    gi::one_electron(C, atom_charges, npc, pc_coords, pc_charges, nullptr, Tm.data(), H0, ldh);
    for (long long j = 0; j < n; ++j)
      for (long long i = 0; i < n; ++i) H0[i + ldh * j] += Tm[i + ldh * j];
  })
}

int gi_chol_accuracy_tau(int64_t accuracy, double *tau) {
  GI_TRY({
    switch (accuracy) {
      case GI_CHOL_ACCURACY_DEFAULT: *tau = 1.0e-5; break;
      case GI_CHOL_ACCURACY_TIGHT: *tau = 1.0e-6; break;
      case GI_CHOL_ACCURACY_LUDICROUS: *tau = 1.0e-7; break;
      case GI_CHOL_ACCURACY_DEBUG: *tau = 1.0e-10; break;
      default: throw Error{6, "Invalid accuracy setting"};
    }
  })
}

int gi_chol_build(gi_context *cp, double tau, double kappa, int64_t max_block_dim, int64_t vec_block,
                  int64_t flags) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (!(tau > 0.0)) throw Error{6, "tau must be positive"};
    gi::chol_build(C, tau, kappa, max_block_dim, vec_block, flags);
  })
}

int gi_chol_dims(gi_context *cp, int64_t dims[8]) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    const gi::CholState &S = C.chol;
    if (!S.valid) throw Error{50, "no Cholesky decomposition is available (run gi_chol_build first)"};
    dims[0] = S.nvecs;
    dims[1] = S.norb_pairs;
    dims[2] = S.nshell_pairs;
    dims[3] = S.nsubsets;
    dims[4] = S.max_subset_dim;
    dims[5] = S.npivot_shell_pairs;
    dims[6] = S.nmacro;
    dims[7] = S.nquartets;
  })
}

static void copy_vec(int64_t *dst, const std::vector<long long> &v) {
  if (dst && !v.empty()) std::memcpy(dst, v.data(), sizeof(int64_t) * v.size());
}

int gi_chol_meta(gi_context *cp, int64_t *shell_pairs, int64_t *shell_pair_loc, int64_t *shell_pair_dim,
                 int64_t *subset_dim, int64_t *subset_bounds, int64_t *pivots, int64_t *pivot_order,
                 int64_t *pivot_shell_pairs, int64_t *pivot_shell_pair_loc, int64_t *pivot_shell_pair_dim,
                 int64_t *pivot_orb_pairs) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    const gi::CholState &S = C.chol;
    if (!S.valid) throw Error{50, "no Cholesky decomposition is available (run gi_chol_build first)"};
    copy_vec(shell_pairs, S.shell_pairs);
    copy_vec(shell_pair_loc, S.shell_pair_loc);
    copy_vec(shell_pair_dim, S.shell_pair_dim);
    copy_vec(subset_dim, S.subset_dim);
    copy_vec(subset_bounds, S.subset_bounds);
    copy_vec(pivots, S.pivots);
    copy_vec(pivot_order, S.pivot_order);
    copy_vec(pivot_shell_pairs, S.pivot_shell_pairs);
    copy_vec(pivot_shell_pair_loc, S.pivot_shell_pair_loc);
    copy_vec(pivot_shell_pair_dim, S.pivot_shell_pair_dim);
    copy_vec(pivot_orb_pairs, S.pivot_orb_pairs);
  })
}

int gi_chol_rkpq(gi_context *cp, double *dst) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    const gi::CholState &S = C.chol;
    if (!S.valid || !S.rkpq.p) throw Error{50, "Cholesky vectors are not available"};
    size_t n = (size_t)S.nvecs * S.max_subset_dim * S.nsubsets;
    GI_HIP(hipMemcpy(dst, S.rkpq.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  })
}

int gi_chol_rkpq_device(gi_context *cp, double **ptr) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (!C.chol.valid || !C.chol.rkpq.p) throw Error{50, "Cholesky vectors are not available"};
    *ptr = C.chol.rkpq.p;
  })
}

int gi_chol_inv_l(gi_context *cp, double *dst) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (!C.chol.valid || !C.chol.inv_l.p) throw Error{50, "Inv_L is not available"};
    GI_HIP(hipMemcpy(dst, C.chol.inv_l.p, sizeof(double) * C.chol.nvecs * C.chol.nvecs, hipMemcpyDeviceToHost));
  })
}

int gi_chol_vdiag(gi_context *cp, double *dst) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (!C.chol.valid || C.chol.vdiag.empty()) throw Error{50, "Vdiag is not available"};
    std::memcpy(dst, C.chol.vdiag.data(), sizeof(double) * C.chol.vdiag.size());
  })
}

int gi_chol_residual(gi_context *cp, double *max_residual, double *min_residual) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (!max_residual || !min_residual) throw Error{8, "null output pointer"};
    gi::chol_residual(C, max_residual, min_residual);
  })
}

int gi_chol_timings(gi_context *cp, double t[18]) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    auto &T = C.chol.timings;
    auto g = [&](const char *k) { auto it = T.find(k); return it == T.end() ? 0.0 : it->second; };
    t[0] = g("step1");
    t[1] = g("step2");
    t[2] = g("vectors");
    t[3] = g("total");
    t[4] = g("eri_kernel");
    t[5] = (double)C.chol.eri_flops;
    t[6] = g("vdiag");
    t[7] = g("eri_step1");
    t[8] = g("rtr");
    t[9] = g("micro");
    t[10] = g("vpqrs");
    t[11] = g("potrf_trtri");
    t[12] = g("vabrs");
    t[13] = g("vectors_gemm");
    t[14] = g("eri_launches");
    t[15] = g("rtr_flops");
    t[16] = g("candidates");
    t[17] = g("batch");
  })
}

int gi_eri_class_stats(gi_context *cp, int64_t cap, double *rows, int64_t *n) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    int64_t k = 0;
    for (auto &kv : C.class_stats) {
      if (k >= cap) break;
      int key = kv.first;
      double *r = rows + 5 * k;
      r[0] = key / 16;
      r[1] = key % 16;
      r[2] = kv.second[0];
      r[3] = kv.second[1];
      r[4] = kv.second[2];
      ++k;
    }
    *n = k;
  })
}

int gi_mo_transform_device(gi_context *cp, const double *CA, int64_t lda, int64_t a0, int64_t a1,
                           const double *CB, int64_t ldb, int64_t b0, int64_t b1, int64_t max_buffer_mb,
                           int64_t flags, double *Rkab) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    gi::mo_transform(C, CA, lda, a0, a1, CB, ldb, b0, b1, max_buffer_mb, flags, Rkab);
  })
}

int gi_mo_transform(gi_context *cp, const double *CA, int64_t lda, int64_t a0, int64_t a1, const double *CB,
                    int64_t ldb, int64_t b0, int64_t b1, int64_t max_buffer_mb, int64_t flags, double *Rkab) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (!C.chol.valid) throw Error{50, "Cholesky vectors are not available (run gi_chol_build first)"};
    int64_t ncolA = a1, ncolB = b1;
    gi::DBuf<double> dA, dB, dR;
    dA.upload(CA, (size_t)lda * ncolA, C.stream);
    dB.upload(CB, (size_t)ldb * ncolB, C.stream);
    size_t nr = (size_t)C.chol.nvecs * (a1 - a0 + 1) * (b1 - b0 + 1);
    dR.alloc(nr);
    gi::mo_transform(C, dA.p, lda, a0, a1, dB.p, ldb, b0, b1, max_buffer_mb, flags, dR.p);
    GI_HIP(hipMemcpy(Rkab, dR.p, sizeof(double) * nr, hipMemcpyDeviceToHost));
  })
}

int gi_ao_map(gi_context *cp, int64_t ordering, int64_t *map, double *phase) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    std::vector<long long> m;
    std::vector<double> ph;
    gi::ao_map(C, (int)ordering, m, ph);
    for (int p = 0; p < C.nao; ++p) {
      map[p] = (int64_t)m[p] + 1;
      phase[p] = ph[p];
    }
  })
}

int gi_ao_rows_to_auto2e_device(gi_context *cp, int64_t ordering, const double *X_ext, int64_t ld_ext,
                                int64_t ncols, double *X_ao, int64_t ld_ao) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (X_ext == X_ao) throw Error{72, "in-place AO reordering is not supported"};
    gi::ao_rows_to_auto2e(C, (int)ordering, X_ext, ld_ext, ncols, X_ao, ld_ao);
  })
}

int gi_ao_rows_to_auto2e_device_noalloc(gi_context *cp, int64_t ordering, const double *X_ext, int64_t ld_ext,
                                        int64_t ncols, double *X_ao, int64_t ld_ao) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (X_ext == X_ao) throw Error{72, "in-place AO reordering is not supported"};
    gi::ao_rows_to_auto2e_noalloc(C, (int)ordering, X_ext, ld_ext, ncols, X_ao, ld_ao);
  })
}

int gi_thc_orbitals_device(gi_context *cp, int64_t npts, const double *xyz, int64_t normalize, double *Xgp,
                           int64_t ldx) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    gi::thc_orbitals(C, xyz, npts, (int)(normalize != 0), Xgp, ldx);
    C.wait();
  })
}

int gi_thc_orbitals(gi_context *cp, int64_t npts, const double *xyz, int64_t normalize, double *Xgp,
                    int64_t ldx) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (npts < 0 || ldx < npts) throw Error{81, "invalid number of grid points or leading dimension"};
    if (npts == 0) return 0;
    gi::DBuf<double> dxyz, dX;
    dxyz.upload(xyz, (size_t)(3 * npts), C.stream);
    dX.alloc((size_t)(ldx * C.nao));
    gi::thc_orbitals(C, dxyz.p, npts, (int)(normalize != 0), dX.p, ldx);
    GI_HIP(hipMemcpy2DAsync(Xgp, sizeof(double) * ldx, dX.p, sizeof(double) * ldx, sizeof(double) * npts, C.nao,
                            hipMemcpyDeviceToHost, C.stream));
    C.wait();
  })
}

int gi_thc_z_device(gi_context *cp, int64_t npts, const double *Xgp, int64_t ldx, int64_t npts_reduced,
                    double *Zgk, int64_t ldz, double *Zgk_reduced, int64_t ldzr) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    gi::thc_z(C, Xgp, ldx, npts, npts_reduced, Zgk, ldz, Zgk_reduced, ldzr);
  })
}

int gi_thc_z(gi_context *cp, int64_t npts, const double *Xgp, int64_t ldx, int64_t npts_reduced, double *Zgk,
             int64_t ldz, double *Zgk_reduced, int64_t ldzr) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (!C.chol.valid) throw Error{50, "Cholesky vectors are not available (run gi_chol_build first)"};
    if (npts <= 0 || ldx < npts || ldz < npts) throw Error{81, "invalid number of grid points or leading dimension"};
    const int64_t nvec = C.chol.nvecs;
    const bool reduced = npts_reduced > 0 && npts_reduced < npts;
    if (reduced && (!Zgk_reduced || ldzr < npts_reduced)) throw Error{81, "invalid ZgkReduced"};
    gi::DBuf<double> dX, dZ, dZr;
    dX.alloc((size_t)(npts * C.nao));
    GI_HIP(hipMemcpy2DAsync(dX.p, sizeof(double) * npts, Xgp, sizeof(double) * ldx, sizeof(double) * npts, C.nao,
                            hipMemcpyHostToDevice, C.stream));
    dZ.alloc((size_t)(npts * nvec));
    if (reduced) dZr.alloc((size_t)(npts_reduced * nvec));
    gi::thc_z(C, dX.p, npts, npts, npts_reduced, dZ.p, npts, dZr.p, npts_reduced);
    GI_HIP(hipMemcpy2DAsync(Zgk, sizeof(double) * ldz, dZ.p, sizeof(double) * npts, sizeof(double) * npts, nvec,
                            hipMemcpyDeviceToHost, C.stream));
    if (reduced)
      GI_HIP(hipMemcpy2DAsync(Zgk_reduced, sizeof(double) * ldzr, dZr.p, sizeof(double) * npts_reduced,
                              sizeof(double) * npts_reduced, nvec, hipMemcpyDeviceToHost, C.stream));
    C.wait();
  })
}

int gi_thc_rkab(gi_context *cp, int64_t npts, const double *Xgp, int64_t ldx, const double *Zgk, int64_t ldz,
                int64_t nvec, const double *CA, int64_t lda, int64_t a0, int64_t a1, const double *CB,
                int64_t ldb, int64_t b0, int64_t b1, int64_t ordering, double *Rkab) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    const int64_t NA = a1 - a0 + 1, NB = b1 - b0 + 1;
    if (NA <= 0 || NB <= 0 || a0 < 1 || b0 < 1) throw Error{51, "Invalid dimension of Rkab"};
    if (npts <= 0 || nvec <= 0 || ldx < npts || ldz < npts) throw Error{81, "invalid Xgp or Zgk"};
    if (lda < C.nao || ldb < C.nao) throw Error{52, "Invalid number of atomic orbitals in CA/CB"};
    gi::DBuf<double> dX, dZ, dA, dB, dR;
    dX.alloc((size_t)(npts * C.nao));
    GI_HIP(hipMemcpy2DAsync(dX.p, sizeof(double) * npts, Xgp, sizeof(double) * ldx, sizeof(double) * npts, C.nao,
                            hipMemcpyHostToDevice, C.stream));
    dZ.alloc((size_t)(npts * nvec));
    GI_HIP(hipMemcpy2DAsync(dZ.p, sizeof(double) * npts, Zgk, sizeof(double) * ldz, sizeof(double) * npts, nvec,
                            hipMemcpyHostToDevice, C.stream));
    dA.upload(CA, (size_t)(lda * a1), C.stream);
    dB.upload(CB, (size_t)(ldb * b1), C.stream);
    const size_t nr = (size_t)(nvec * NA * NB);
    dR.alloc(nr);
    gi::thc_rkab(C, dX.p, npts, npts, dZ.p, npts, nvec, dA.p, lda, a0, a1, dB.p, ldb, b0, b1, (int)ordering, dR.p);
    GI_HIP(hipMemcpyAsync(Rkab, dR.p, sizeof(double) * nr, hipMemcpyDeviceToHost, C.stream));
    C.wait();
  })
}

int gi_thc_rkab_device(gi_context *cp, int64_t npts, const double *Xgp, int64_t ldx, const double *Zgk,
                       int64_t ldz, int64_t nvec, const double *CA, int64_t lda, int64_t a0, int64_t a1,
                       const double *CB, int64_t ldb, int64_t b0, int64_t b1, int64_t ordering, double *Rkab) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    gi::thc_rkab(C, Xgp, ldx, npts, Zgk, ldz, nvec, CA, lda, a0, a1, CB, ldb, b0, b1, (int)ordering, Rkab);
  })
}

int gi_thc_grid_pivots(gi_context *cp, int64_t npts, const double *xyz, double qr_thresh,
                       double qr_thresh_reduced, int64_t block_dim, int64_t *npivots, int64_t *npivots_reduced,
                       int64_t *pivots) {
  GI_TRY({
    gi::Context &C = ctx_of(cp);
    if (npts <= 0) throw Error{81, "invalid number of grid points"};
    if (!npivots || !npivots_reduced || !pivots) throw Error{7, "null output pointer"};
    gi::DBuf<double> dxyz;
    dxyz.upload(xyz, (size_t)(3 * npts), C.stream);
    std::vector<long long> piv;
    long long nred = 0;
    gi::thc_grid_pivots(C, dxyz.p, npts, qr_thresh, qr_thresh_reduced, block_dim, piv, nred);
    for (size_t i = 0; i < piv.size(); ++i) pivots[i] = (int64_t)piv[i];
    *npivots = (int64_t)piv.size();
    *npivots_reduced = (int64_t)nred;
    C.wait();
  })
}

}  // extern "C"
