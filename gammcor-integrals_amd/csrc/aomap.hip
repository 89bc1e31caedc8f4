// External-program AO orderings: the permutation from an external program's
// order of the angular functions to the Auto2e order used by every kernel of
// the library, the Orca sign convention, and the device kernel that applies
// both to the rows of MO coefficient matrices, so that coefficients in an
// external order go straight to gi_mo_transform_device.
//
// Conventions (auto2e_interface_AOMap, Auto2eInterface.f90:378-514): within a
// shell the Auto2e order of spherical functions is m = -l..l for l >= 2 and
// (x, y, z) for p; Cartesian functions run lx = l..0, ly = l - lx..0.  The
// shells themselves are in the same order in both programs, except for Open
// Molcas, which orders the AOs of each atom by l, then m, then shell.
#include <mutex>
#include "gi_context.h"

namespace gi {

namespace {

constexpr int AO_LMAX = 5;
constexpr int ORD_AUTO2E = 0, ORD_MOLPRO = 1, ORD_ORCA = 2, ORD_DALTON = 3, ORD_MOLCAS = 4, ORD_PYSCF = 5;

// Auto2e position (0-based, within the shell) of the external program's
// i-th spherical function, per l: the tables of Auto2eInterface.f90:24-83
// written as offsets into m = -l..l (p: x, y, z).
const int *spher_table(int ordering, int l) {
  static const int ident[6][11] = {{0}, {0, 1, 2}, {0, 1, 2, 3, 4}, {0, 1, 2, 3, 4, 5, 6},
                                   {0, 1, 2, 3, 4, 5, 6, 7, 8}, {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10}};
  // Molpro: d0 d2- d1+ d2+ d1- / f1+ f1- f0 f3+ f2- f3- f2+ / ...
  static const int molpro[6][11] = {{0},
                                    {0, 1, 2},
                                    {2, 0, 3, 4, 1},
                                    {4, 2, 3, 6, 1, 0, 5},
                                    {4, 2, 5, 8, 3, 6, 0, 7, 1},
                                    {6, 4, 7, 8, 1, 2, 9, 0, 5, 10, 3}};
  // Orca: p as (z, x, y); l >= 2 as m = 0, +1, -1, +2, -2, ...
  static const int orca[6][11] = {{0},
                                  {2, 0, 1},
                                  {2, 3, 1, 4, 0},
                                  {3, 4, 2, 5, 1, 6, 0},
                                  {4, 5, 3, 6, 2, 7, 1, 8, 0},
                                  {5, 6, 4, 7, 3, 8, 2, 9, 1, 10, 0}};
  switch (ordering) {
    case ORD_MOLPRO: return molpro[l];
    case ORD_ORCA: return orca[l];
    case ORD_DALTON:
    case ORD_MOLCAS:
    case ORD_PYSCF: return ident[l];
    default: throw Error{70, "Invalid external orbital ordering"};
  }
}

// Molpro's Cartesian functions (Auto2eInterface.f90:150-213): xx yy zz xy xz
// yz for d, and so on; h is in the Auto2e order.
const int *cart_table_molpro(int l) {
  static const int t[6][21] = {{0},
                               {0, 1, 2},
                               {0, 3, 5, 1, 2, 4},
                               {0, 6, 9, 1, 2, 3, 7, 5, 8, 4},
                               {0, 10, 14, 1, 2, 6, 11, 9, 13, 3, 5, 12, 4, 7, 8},
                               {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20}};
  return t[l];
}

}  // namespace

void ao_map(const Context &C, int ordering, std::vector<long long> &map, std::vector<double> &phase) {
  if (!C.basis_set) throw Error{4, "basis not set"};
  const int nsh = C.nshells;
  for (int s = 0; s < nsh; ++s)
    if (C.shell_l[s] > AO_LMAX) throw Error{70, "angular momentum beyond the AO ordering tables"};
  map.assign(C.nao, -1);
  phase.assign(C.nao, 1.0);
  if (C.spher) {
    if (ordering == ORD_MOLCAS) {
      // per atom: l, then m, then the atom's shells of that l in order
      int s0 = 0;
      while (s0 < nsh) {
        int s1 = s0;
        while (s1 + 1 < nsh && C.shell_atom[s1 + 1] == C.shell_atom[s0]) ++s1;
        int lmax = 0;
        for (int s = s0; s <= s1; ++s) lmax = std::max(lmax, C.shell_l[s]);
        long long i = C.shell_loc[s0];
        for (int l = 0; l <= lmax; ++l)
          for (int m = 0; m < 2 * l + 1; ++m)
            for (int s = s0; s <= s1; ++s)
              if (C.shell_l[s] == l) map[i++] = C.shell_loc[s] + spher_table(ordering, l)[m];
        s0 = s1 + 1;
      }
    } else {
      for (int s = 0; s < nsh; ++s) {
        const int *t = spher_table(ordering, C.shell_l[s]);
        for (int i = 0; i < C.shell_nang[s]; ++i) map[C.shell_loc[s] + i] = C.shell_loc[s] + t[i];
      }
    }
    if (ordering == ORD_ORCA) {
      // Orca's signs of the real combinations m = +-3 (f, g, h) and m = +-4
      // (g, h) (auto2e_interface_ApplyOrcaPhases_Vector, Auto2eInterface.f90:517-562)
      for (int s = 0; s < nsh; ++s) {
        const int l = C.shell_l[s];
        const long long c = C.shell_loc[s] + l;   // Auto2e position of m = 0
        if (l >= 3) phase[c + 3] = phase[c - 3] = -1.0;
        if (l >= 4) phase[c + 4] = phase[c - 4] = -1.0;
      }
    }
  } else {
    if (ordering != ORD_MOLPRO) throw Error{70, "Invalid external orbital ordering"};
    for (int s = 0; s < nsh; ++s) {
      const int *t = cart_table_molpro(C.shell_l[s]);
      for (int i = 0; i < C.shell_nang[s]; ++i) map[C.shell_loc[s] + i] = C.shell_loc[s] + t[i];
    }
  }
  std::vector<char> seen(C.nao, 0);
  for (long long p = 0; p < C.nao; ++p) {
    if (map[p] < 0 || map[p] >= C.nao || seen[map[p]]) throw Error{71, "AO ordering map is not a permutation"};
    seen[map[p]] = 1;
  }
}

// X_ao(map[p], j) = phase[map[p]] X_ext(p, j), p < nao, j < ncols
__global__ void k_ao_rows(const double *__restrict__ X, long long ldx, long long nao, long long ncols,
                          const long long *__restrict__ map, const double *__restrict__ phase,
                          double *__restrict__ Y, long long ldy) {
  const long long total = nao * ncols;
  for (long long e = threadIdx.x + (long long)blockIdx.x * blockDim.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long p = e % nao, j = e / nao;
    const long long q = map[p];
    Y[q + ldy * j] = phase[q] * X[p + ldx * j];
  }
}

void ao_rows_to_auto2e(Context &C, int ordering, const double *X, long long ldx, long long ncols, double *Y,
                       long long ldy) {
  std::vector<long long> map;
  std::vector<double> phase;
  ao_map(C, ordering, map, phase);
  if (ldx < C.nao || ldy < C.nao) throw Error{52, "Invalid number of atomic orbitals in CA/CB"};
  if (ncols <= 0) return;
  DBuf<long long> dmap;
  DBuf<double> dph;
  dmap.upload(map, C.stream);
  dph.upload(phase, C.stream);
  const long long total = (long long)C.nao * ncols;
  const unsigned grid = (unsigned)std::min<long long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_ao_rows, dim3(grid), dim3(256), 0, C.stream, X, ldx, (long long)C.nao, ncols, dmap.p, dph.p,
                     Y, ldy);
  GI_HIP(hipGetLastError());
  GI_HIP(hipStreamSynchronize(C.stream));
}

// ---- allocation-free variant (gi_ao_rows_to_auto2e_device_noalloc)
// The map and the signs are staged in a file-scope device array instead of
// stream-ordered allocations: NOALLOC_MAX_NAO map entries (int64) followed by
// NOALLOC_MAX_NAO signs (double), 1 MiB, held on every device the library
// runs on until the process ends.  Host code never passes `scratch` itself
// to a kernel or copy (in host code the name is a stand-in, not a device
// address): the device address comes from hipGetSymbolAddress, taken once per
// device with that device current (noalloc_setup, called by gi_create), and
// only that address is used.  Every call on one device shares the array, so
// overlapping calls on one device race; each call writes the whole map and
// sign range it reads, so nothing has to start at zero.
constexpr long long NOALLOC_MAX_NAO = 1LL << 16;
static __device__ __align__(256) unsigned char scratch[NOALLOC_MAX_NAO * (sizeof(long long) + sizeof(double))];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

// with `device` current: fetch the device address of `scratch` once per device
void noalloc_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) throw Error{1, "device index beyond the library's table"};
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  GI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Error{2, "hipGetSymbolAddress returned no address for the scratch array"};
  g_scratch_ptr[device] = p;
}

void ao_rows_to_auto2e_noalloc(Context &C, int ordering, const double *X, long long ldx, long long ncols, double *Y,
                               long long ldy) {
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (C.device >= 0 && C.device < NOALLOC_MAX_DEVICES) base = g_scratch_ptr[C.device];
  }
  if (!base) throw Error{74, "scratch array of the allocation-free entry points is not set up on this device"};
  if (C.nao > NOALLOC_MAX_NAO)
    throw Error{73, "allocation-free AO reordering supports at most " + std::to_string(NOALLOC_MAX_NAO) + " AOs"};
  std::vector<long long> map;
  std::vector<double> phase;
  ao_map(C, ordering, map, phase);
  if (ldx < C.nao || ldy < C.nao) throw Error{52, "Invalid number of atomic orbitals in CA/CB"};
  if (ncols <= 0) return;
  long long *dmap = static_cast<long long *>(base);
  double *dph = reinterpret_cast<double *>(static_cast<unsigned char *>(base) + NOALLOC_MAX_NAO * sizeof(long long));
  GI_HIP(hipMemcpyAsync(dmap, map.data(), sizeof(long long) * C.nao, hipMemcpyHostToDevice, C.stream));
  GI_HIP(hipMemcpyAsync(dph, phase.data(), sizeof(double) * C.nao, hipMemcpyHostToDevice, C.stream));
  const long long total = (long long)C.nao * ncols;
  const unsigned grid = (unsigned)std::min<long long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_ao_rows, dim3(grid), dim3(256), 0, C.stream, X, ldx, (long long)C.nao, ncols, dmap, dph, Y,
                     ldy);
  GI_HIP(hipGetLastError());
  GI_HIP(hipStreamSynchronize(C.stream));
}

}  // namespace gi
