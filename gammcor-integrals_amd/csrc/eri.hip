// ERI launch layer: instantiates one kernel per canonical class (list emitted
// by gigpu/eri_gen.py into eri_classes.inc), bins shell pairs by angular
// momenta and primitive counts, and launches the Cartesian product (or the
// zipped pairing) of every row bin with every column bin.
#include <algorithm>
#include <cstdio>
#include <map>
#include <tuple>
#include "gi_context.h"

namespace gi {

__constant__ double c_spher[SPHER_TOTAL];

using LaunchFn = void (*)(const EriArgs &, long long ngroups, hipStream_t);

template <int LA, int LB, int LC, int LD>
static void launch_class(const EriArgs &a, long long ngroups, hipStream_t s) {
  using P = EriPlan<LA, LB, LC, LD>;
  long long grid = ngroups;
  if (P::MODE == MODE_GLB && grid > 2048) grid = 2048;
  if (grid > 0x7fffffffLL) grid = 0x7fffffffLL;
  hipLaunchKernelGGL((eri_class_kernel<LA, LB, LC, LD>), dim3((unsigned)grid), dim3(64), 0, s, a);
}

struct ClassEntry {
  LaunchFn fn;
  ClassPlan plan;
};

static std::map<int, ClassEntry> &class_table() {
  static std::map<int, ClassEntry> t;
  if (t.empty()) {
#define GI_CLASS(a, b, c, d, fl)                                                               \
  t[((a)*8 + (b)) * 64 + (c)*8 + (d)] = ClassEntry{                                           \
      &launch_class<a, b, c, d>,                                                              \
      ClassPlan{EriPlan<a, b, c, d>::MODE, EriPlan<a, b, c, d>::G, EriPlan<a, b, c, d>::QPW, \
                EriPlan<a, b, c, d>::PER_Q, (long long)(fl)}};
#include "eri_classes.inc"
#undef GI_CLASS
  }
  return t;
}

const ClassPlan *class_plan(int la, int lb, int lc, int ld) {
  auto &t = class_table();
  auto it = t.find((la * 8 + lb) * 64 + lc * 8 + ld);
  return it == t.end() ? nullptr : &it->second.plan;
}

bool class_supported(int la, int lb, int lc, int ld) { return class_plan(la, lb, lc, ld) != nullptr; }

void upload_constants(Context &C) {
  std::vector<double> U(SPHER_TOTAL, 0.0);
  for (int l = 2; l <= MAXL; ++l) spher_u_host(l, U.data() + spher_off(l));
  GI_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_spher), U.data(), sizeof(double) * SPHER_TOTAL));
  std::vector<double> table;
  boys_table_host(BOYS_MAXM, table);
  C.dboys.upload(table, C.stream);
  GI_HIP(hipStreamSynchronize(C.stream));
}

// canonical order of (la, lb, lc, ld); m[k] = caller slot of canonical position k
static void canonicalize(const int l_in[4], int lc_out[4], int m[4]) {
  int l[4] = {l_in[0], l_in[1], l_in[2], l_in[3]};
  for (int k = 0; k < 4; ++k) m[k] = k;
  if (l[1] > l[0]) { std::swap(l[0], l[1]); std::swap(m[0], m[1]); }
  if (l[3] > l[2]) { std::swap(l[2], l[3]); std::swap(m[2], m[3]); }
  if (l[2] > l[0] || (l[2] == l[0] && l[3] > l[1])) {
    std::swap(l[0], l[2]); std::swap(l[1], l[3]);
    std::swap(m[0], m[2]); std::swap(m[1], m[3]);
  }
  for (int k = 0; k < 4; ++k) lc_out[k] = l[k];
}

void eri_run(Context &C, const EriRequest &R) {
  auto &tab = class_table();
  // bin key of a pair: (l0, l1, np0, np1)
  auto key_of = [&](const PairDesc &d) {
    return std::make_tuple(C.shell_l[d.s0], C.shell_l[d.s1], C.shell_np[d.s0], C.shell_np[d.s1]);
  };
  using Key = std::tuple<int, int, int, int>;
  std::map<Key, std::vector<int>> rbins, cbins;
  for (int r : *R.rows) rbins[key_of((*R.rdesc)[r])].push_back(r);
  if (R.mode == LAUNCH_PRODUCT) {
    for (int c : *R.cols) cbins[key_of((*R.cdesc)[c])].push_back(c);
  }
  // For zipped launches the row and column lists are paired element by
  // element; bin by the combined key.
  struct Launch {
    Key rk, ck;
    std::vector<int> rows, cols;
  };
  std::vector<Launch> launches;
  if (R.mode == LAUNCH_PRODUCT) {
    for (auto &rb : rbins)
      for (auto &cb : cbins) launches.push_back(Launch{rb.first, cb.first, rb.second, cb.second});
  } else {
    std::map<std::pair<Key, Key>, Launch> zb;
    for (size_t i = 0; i < R.rows->size(); ++i) {
      int r = (*R.rows)[i], c = (*R.cols)[i];
      Key rk = key_of((*R.rdesc)[r]), ck = key_of((*R.cdesc)[c]);
      auto &L = zb[{rk, ck}];
      L.rk = rk;
      L.ck = ck;
      L.rows.push_back(r);
      L.cols.push_back(c);
    }
    for (auto &kv : zb) launches.push_back(std::move(kv.second));
  }
  // one device index buffer for all launches
  std::vector<int> idx;
  std::vector<size_t> roff(launches.size()), coff(launches.size());
  for (size_t i = 0; i < launches.size(); ++i) {
    roff[i] = idx.size();
    idx.insert(idx.end(), launches[i].rows.begin(), launches[i].rows.end());
    coff[i] = idx.size();
    idx.insert(idx.end(), launches[i].cols.begin(), launches[i].cols.end());
  }
  DBuf<int> didx;
  didx.upload(idx, C.stream);
  // global workspace for the largest classes
  size_t need = 0;
  for (auto &L : launches) {
    int l[4] = {std::get<0>(L.rk), std::get<1>(L.rk), std::get<0>(L.ck), std::get<1>(L.ck)};
    int lc[4], m[4];
    canonicalize(l, lc, m);
    const ClassPlan *pl = class_plan(lc[0], lc[1], lc[2], lc[3]);
    if (!pl) {
      char buf[160];
      snprintf(buf, sizeof buf, "no ERI kernel for class (%d%d|%d%d); shells with l > 3 are not supported",
               lc[0], lc[1], lc[2], lc[3]);
      throw Error{3, buf};
    }
    if (pl->mode == MODE_GLB) need = std::max(need, (size_t)2048 * pl->per_q);
  }
  if (need) C.gws.alloc(need);

  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (C.time_eri) {
    GI_HIP(hipEventCreate(&e0));
    GI_HIP(hipEventCreate(&e1));
    GI_HIP(hipEventRecord(e0, C.stream));
  }
  long long nq_total = 0, flops_total = 0;
  static const bool stats = getenv("GIGPU_ERI_STATS") != nullptr;
  for (size_t i = 0; i < launches.size(); ++i) {
    auto &L = launches[i];
    int l[4] = {std::get<0>(L.rk), std::get<1>(L.rk), std::get<0>(L.ck), std::get<1>(L.ck)};
    int np[4] = {std::get<2>(L.rk), std::get<3>(L.rk), std::get<2>(L.ck), std::get<3>(L.ck)};
    int lc[4], m[4];
    canonicalize(l, lc, m);
    auto &ent = tab.at((lc[0] * 8 + lc[1]) * 64 + lc[2] * 8 + lc[3]);
    EriArgs a{};
    a.shells = C.dshells.p;
    a.exps = C.dexps.p;
    a.cntr = C.dcntr.p;
    a.norms = C.dnorms.p;
    a.boys = C.dboys.p;
    a.rdesc = R.d_rdesc;
    a.cdesc = R.d_cdesc;
    a.rows = didx.p + roff[i];
    a.cols = didx.p + coff[i];
    a.nrows = (long long)L.rows.size();
    a.ncols = (long long)L.cols.size();
    a.mode = R.mode;
    a.perm = m[0] | (m[1] << 2) | (m[2] << 4) | (m[3] << 6);
    a.spher = C.spher;
    a.kappa = R.kappa;
    a.out = R.out;
    a.ldo = R.ldo;
    a.sel = R.d_sel;
    a.gws = C.gws.p;
    a.npa = np[m[0]];
    a.npb = np[m[1]];
    a.npc = np[m[2]];
    a.npd = np[m[3]];
    long long ntask = (R.mode == LAUNCH_PRODUCT) ? a.nrows * a.ncols : a.nrows;
    if (ntask == 0) continue;
    long long ngroups = (ntask + ent.plan.qpw - 1) / ent.plan.qpw;
    hipEvent_t s0 = nullptr, s1 = nullptr;
    if (stats) {
      GI_HIP(hipEventCreate(&s0));
      GI_HIP(hipEventCreate(&s1));
      GI_HIP(hipEventRecord(s0, C.stream));
    }
    ent.fn(a, ngroups, C.stream);
    GI_HIP(hipGetLastError());
    if (stats) {
      GI_HIP(hipEventRecord(s1, C.stream));
      GI_HIP(hipEventSynchronize(s1));
      float ms = 0;
      GI_HIP(hipEventElapsedTime(&ms, s0, s1));
      auto &st = C.class_stats[(lc[0] * 8 + lc[1]) * 64 + lc[2] * 8 + lc[3]];
      st[0] += ntask;
      st[1] += ms;
      st[2] += (double)ntask * ent.plan.flops * (np[0] * np[1] * np[2] * np[3]);
      (void)hipEventDestroy(s0);
      (void)hipEventDestroy(s1);
    }
    nq_total += ntask;
    flops_total += ntask * ent.plan.flops * (long long)(np[0] * np[1] * np[2] * np[3]);
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
  C.eri_flops_acc += flops_total;
  // didx must outlive the kernels
  GI_HIP(hipStreamSynchronize(C.stream));
}

}  // namespace gi
