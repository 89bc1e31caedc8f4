// Library context: one per (process, device).  Owns the device copy of the
// basis, the Boys table, the ERI workspace and the Cholesky results.
#pragma once
#include <rocblas/rocblas.h>
#include <algorithm>
#include <array>
#include <map>
#include <unordered_map>
#include <memory>
#include <vector>
#include "md_kernel.h"
#include "md4_kernel.h"
#include "gi_common.h"

namespace gi {

// Stream of the context the calling thread works for (set on every entry
// through the C ABI).  Device buffers are allocated and freed stream-ordered
// on it (hipMallocAsync / hipFreeAsync), so contexts driven from different
// host threads never synchronize each other through allocation.
extern thread_local hipStream_t g_alloc_stream;

// owning device buffer
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  hipStream_t st = nullptr;   // stream the buffer was allocated on
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFreeAsync(p, st);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) count = 1;
    st = g_alloc_stream;
    GI_HIP(hipMallocAsync((void **)&p, count * sizeof(T), st));
    n = count;
  }
  void upload(const T *src, size_t count, hipStream_t s) {
    alloc(count);
    if (count) GI_HIP(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
  // grow to at least `count` elements keeping the first n (stream-ordered copy)
  void grow_keep(size_t count, hipStream_t s) {
    if (count <= n && p) return;
    size_t nn = std::max(count, n + n / 2);
    T *q = nullptr;
    GI_HIP(hipMallocAsync((void **)&q, nn * sizeof(T), s));
    if (p && n) GI_HIP(hipMemcpyAsync(q, p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    release();
    p = q;
    n = nn;
    st = s;
  }
};

constexpr int MD_EMAX = 6;   // largest l0 + l1 of a shell pair with an ERI kernel

// Cholesky outputs kept on the device / host (TChol2Vecs + Rkpq)
struct CholState {
  bool valid = false;
  double tau = 0, kappa = 0;
  long long nvecs = 0, norb_pairs = 0, nshell_pairs = 0, nsubsets = 0, max_subset_dim = 0;
  long long npivot_shell_pairs = 0, nmacro = 0, nquartets = 0;
  double max_residual = 0;
  std::vector<long long> shell_pairs, shell_pair_loc, shell_pair_dim, subset_dim, subset_bounds,
      pivots, pivot_order, pivot_shell_pairs, pivot_shell_pair_loc, pivot_shell_pair_dim,
      pivot_orb_pairs;
  std::vector<double> vdiag;
  DBuf<double> inv_l;   // nvecs x nvecs
  DBuf<double> rkpq;    // nvecs x max_subset_dim x nsubsets
  // MO transform pair lists: for AO q, entries mo_adj_off[q]..mo_adj_off[q+1]-1
  // hold the partner AO p and the Rkpq column of the orbital pair {p, q}
  DBuf<long long> mo_adj_off, mo_adj_col;
  DBuf<int> mo_adj_p;
  std::map<std::string, double> timings;
  std::map<std::string, double> eri_kernel_ms;   // per phase ERI kernel time
  std::map<std::string, long long> eri_quartets; // per phase quartet counts
  long long eri_flops = 0;
};

struct CholWork;   // gi_chol_build work arrays (chol.hip)
struct ThcWork;    // work arrays of the THC calls (thc.hip)

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  // basis
  int natoms = 0, nshells = 0, spher = 1, lmax = 0, nao = 0;
  std::vector<DevShell> hshells;
  std::vector<int> shell_l, shell_nang, shell_np, shell_loc;   // shell_loc: 0-based first AO
  std::vector<int> shell_atom;                                  // 0-based atom of each shell
  std::vector<double> atom_xyz;                                 // (3, natoms) bohr
  DBuf<DevShell> dshells;
  DBuf<double> dexps, dcntr, dnorms, dboys;
  bool basis_set = false;
  // ERI pair tables, cached per basis: shell pair (s0, s1) -> entry of
  // md_hdr_host; PrimRec and ET blocks on the device grow as pairs are added
  std::unordered_map<long long, int> md_pt;
  std::vector<int> md_pt_dense;   // s0 * nshells + s1 -> entry or -1 (bases up to 4096 shells)
  std::vector<PairHdr> md_hdr_host;
  long long md_et_used = 0, md_prim_used = 0;
  DBuf<PairHdr> md_hdr;
  DBuf<PrimRec> md_prim;
  DBuf<double> md_et;
  // per request: one device block with the pair-table jobs, descriptors,
  // launch index lists and 4x4x4 column tiles (copied from a pinned host
  // staging buffer), and the packed pair records
  DBuf<unsigned char> md_blob;
  DBuf<QRec> md_rec;
  unsigned char *md_stage = nullptr;   // pinned (hipHostMalloc), grow-only
  size_t md_stage_cap = 0;
  hipEvent_t md_stage_ev = nullptr;    // recorded after each copy out of md_stage
  // pinned staging of at least n bytes, once the last copy out of it is done
  unsigned char *stage_host(size_t n);
  void reset_pair_cache() {
    md_pt.clear();
    if (nshells > 0 && nshells <= 4096)
      md_pt_dense.assign((size_t)nshells * (size_t)nshells, -1);
    else
      md_pt_dense.clear();
    md_hdr_host.clear();
    md_et_used = md_prim_used = 0;
  }
  CholState chol;
  // device work arrays of gi_chol_build (chol.hip), grow-only across builds
  CholWork *cw = nullptr;
  // device work arrays of the THC calls (thc.hip), grow-only
  ThcWork *thcw = nullptr;
  // split-BF16 products (split.hip): BF16 slice planes and per-chunk row
  // scale exponents of the two operands; grow-only, kept across calls
  DBuf<unsigned short> spl_slices[2];
  DBuf<int> spl_exp[2];
  // profiling of ERI launches (events on the library stream)
  bool time_eri = false;
  double eri_ms_acc = 0;
  long long eri_quartets_acc = 0;
  long long eri_flops_acc = 0;
  long long eri_launches_acc = 0;   // kernel dispatches (pair tables + quartet kernels)
  // per kernel class (key 16*E + F, E = la+lb, F = lc+ld): quartets, kernel ms, FP64 ops
  // (filled when GIGPU_ERI_STATS is set in the environment)
  std::map<int, std::array<double, 3>> class_stats;
  // launch-layer tuning (gi_set_option)
  long long eri_min_tiles = 8192, eri_max_kt = 32;
  bool eri_lane = true, eri_trans = true;
  bool eri_md4 = true;   // products on the 4x4x4 kernel (md4_kernel.h) instead of md_kernel
  // products with an s-s side (E = 0 or F = 0) whose other side has order >= this
  // run on the thin-ket form of the 4x4x4 kernel (7: never); below 2 the lane
  // kernels are faster (measured on the (H2O)16 decomposition)
  long long eri_thin_emin = 2;
  // 4x4x4 launches take four bra pairs per tile, or two or one when four would
  // leave fewer tiles than this
  long long eri_md4_tiles = 2048;
  long long eri_lane_emax = LANE_EMAX;   // lane kernels only for classes with E, F <= this
  long long vec_blocks = 20;             // block rows of the triangular Inv_L Vabrs^T product
  // micro iterations of a macro iteration in one launch (one workgroup) when
  // the batch block has at most this many rows
  long long micro_persist_max = 512;
  // Coulomb energy / gradient: shell pairs whose density-weighted Schwarz
  // bound on every term is below 10^-grad_screen are left out (0: none)
  long long grad_screen = 0;
  // THC grid pruning: the explicit Schur-complement form is used once the
  // candidates' ncand^2 doubles fit this many MiB (0: never)
  long long thc_schur_mb = 24576;
  // THC least-squares Z: Cholesky of the metric first (LU when a pivot is not
  // positive); 0 always solves by LU
  int thc_cholesky = 1;
  // class launches of one ERI request are spread over this many streams (the
  // context stream and eri_streams - 1 side streams forked from it and joined
  // back before the request returns); their destination blocks are disjoint
  // host waits for the context stream: hipStreamSynchronize (a spin wait,
  // the lowest latency) by default; with blocking_waits (the batched form,
  // where many host threads wait at once) an event that sleeps until the GPU
  // signals, so that waiting threads leave the host cores to the others
  bool blocking_waits = false;
  hipEvent_t wait_ev = nullptr;
  void wait();
  // small host <-> device transfers of the build through a pinned arena
  // (grow-only): h2d stages the host data there and copies it on the context
  // stream (the arena is reused once wait() has seen every copy complete);
  // read copies device data to the host synchronously
  unsigned char *xfer = nullptr;
  size_t xfer_cap = 0, xfer_used = 0;
  void h2d(void *dst, const void *src, size_t bytes);
  void read(void *dst, const void *src, size_t bytes);
  unsigned char *xfer_slot(size_t bytes);
  static constexpr int ERI_MAX_STREAMS = 4;
  long long eri_streams = ERI_MAX_STREAMS;
  hipStream_t eri_side[ERI_MAX_STREAMS - 1] = {};
  hipEvent_t eri_fork = nullptr, eri_join[ERI_MAX_STREAMS - 1] = {};

  Context(int dev);
  ~Context();
};

// ---------------------------------------------------------------- ERI driver
// Evaluate the quartets of a row-list x column-list product (or zipped lists)
// into a column-major destination, binning pairs by l0 + l1.
struct EriRequest {
  const std::vector<PairDesc> *rdesc = nullptr;   // host lists
  const std::vector<PairDesc> *cdesc = nullptr;
  const std::vector<int> *rows = nullptr, *cols = nullptr;
  int mode = LAUNCH_PRODUCT;
  double *out = nullptr;
  long long ldo = 0;
  const int *d_sel = nullptr;
  double kappa = 0;
};
void eri_run(Context &C, const EriRequest &R);

// one-time device tables
void upload_constants(Context &C);

// C = beta C + alpha op(A) op(B) from split-BF16 products with FP32
// accumulation (split.hip)
void split_bf16_gemm(Context &C, rocblas_operation ta, rocblas_operation tb, long long m, long long n, long long k,
                     double alpha, const double *A, long long lda, const double *B, long long ldb, double beta,
                     double *Cd, long long ldc);

// external AO order -> Auto2e order (aomap.hip): map[p] = Auto2e position of
// external AO p (0-based), phase in Auto2e order (Orca signs, else 1)
void ao_map(const Context &C, int ordering, std::vector<long long> &map, std::vector<double> &phase);
// rows of X (nao x ncols, device) to Auto2e order with the phases applied
void ao_rows_to_auto2e(Context &C, int ordering, const double *X, long long ldx, long long ncols, double *Y,
                       long long ldy);
// allocation-free variant: map and signs staged in a static device array
void ao_rows_to_auto2e_noalloc(Context &C, int ordering, const double *X, long long ldx, long long ncols, double *Y,
                               long long ldy);
void noalloc_setup(int device);   // with `device` current; once per device

// frees the context's gi_chol_build work arrays (chol.hip)
void release_chol_work(Context &C);

// max and min over BASE orbital pairs of (pq|pq) - sum_k Rkpq(k, pq)^2 (chol.hip)
void chol_residual(Context &C, double *rmax, double *rmin);
// Cholesky factor and inverse of a symmetric positive definite n x n matrix
// (chol.hip): L over V's lower triangle, L^-1 into X; false when a pivot is
// not positive.  W holds chol_factor_inverse_scratch(n) doubles.
bool chol_factor_inverse(Context &C, double *V, double *X, long long n, double *W, int *info);
long long chol_factor_inverse_scratch(long long n);

// one md_kernel<E, F> launch (eri.hip)
void launch_md_class(int E, int F, const MdArgs &a, int grid, hipStream_t s);

// density-contracted pair tables of the Coulomb gradient (grad_kernels.h, eri.hip)
constexpr int GRAD_NCOL = 10;   // columns per primitive pair table
struct GradJob {
  int s0, s1, ia, ib;   // shells (s0 >= s1) and primitive indices
  int prim, pad;        // PrimRec slot
  long long et;         // table offset (doubles)
  double w;             // density weight: 2 for two different shells, else 1
};
void launch_build_grad_et(const Context &C, const int *ao_loc, const double *D, long long ldd,
                          const GradJob *jobs, long long njobs, PrimRec *prims, double *et);

// one-electron integrals (onee_kernels.h with eri.hip, onee.hip)
constexpr int ONEE_LMAX_HOST = 3;
void launch_one_electron(const Context &C, int L, const int *ao_loc, const int *pairs, long long npairs,
                         const double *chg, int ncharge, double *S, double *T, double *V, long long ldh);
void one_electron(Context &C, const double *atom_charges, long long npc, const double *pc_xyz, const double *pc_q,
                  double *S, double *T, double *V, long long ldh);

// tensor hypercontraction (thc_kernels.h with eri.hip, thc.hip); device pointers
void launch_thc_orbitals(const Context &C, int L, const int *ao_loc, const int *shl, long long ns,
                         const double *xyz, long long npts, double *X, long long ldx);
void thc_orbitals(Context &C, const double *dxyz, long long npts, int normalize, double *dX, long long ldx);
void thc_z(Context &C, const double *dX, long long ldx, long long npts, long long nred, double *dZ, long long ldz,
           double *dZr, long long ldzr);
void thc_rkab(Context &C, const double *dX, long long ldx, long long npts, const double *dZ, long long ldz,
              long long nvec, const double *dCA, long long lda, long long a0, long long a1, const double *dCB,
              long long ldb, long long b0, long long b1, int ordering, double *dRkab);
void thc_grid_pivots(Context &C, const double *dxyz, long long npts, double qr_thresh, double qr_thresh_reduced,
                     long long block_dim, std::vector<long long> &pivots, long long &nred);
void release_thc_work(Context &C);

// Coulomb energy of a density and its nuclear gradient (grad.hip)
void coulomb_energy_grad(Context &C, const double *D, long long ldd, double kappa, double *energy,
                         double *grad);

}  // namespace gi
