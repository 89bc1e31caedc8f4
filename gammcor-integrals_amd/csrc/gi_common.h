// Shared definitions of the MI355X Cholesky/ERI library (libgigpu).
// Host + device.  No torch types anywhere: the C ABI in include/gigpu.h is the
// only boundary.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

namespace gi {

// ---------------------------------------------------------------- errors
struct Error {
  int code;
  std::string msg;
};
void set_error(const std::string &m);
const char *last_error();

#define GI_HIP(call)                                                             \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      throw gi::Error{2, std::string("HIP error ") + hipGetErrorString(e_) +     \
                             " at " + __FILE__ + ":" + std::to_string(__LINE__)}; \
  } while (0)

// ---------------------------------------------------------------- index algebra
// Cartesian components of a shell are ordered lx = l..0, ly = l-lx..0
// (Auto2eGlobal.py:35-43).  Hermite triples (t,u,v) are packed by total order
// n = t+u+v, and inside one n in the same lx-major order, so the packing of a
// lower total order is a prefix of a higher one.
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
__host__ __device__ constexpr int cidx(int lx, int ly, int lz) {
  return (lx + ly + lz - lx) * (lx + ly + lz - lx + 1) / 2 + (lz);
}
__host__ __device__ constexpr int hidx(int t, int u, int v) {
  return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + cidx(t, u, v);
}
// inverse of cidx inside shell l: returns packed lx | ly<<8 | lz<<16
__host__ __device__ constexpr int cart_of(int l, int i) {
  int s = 0;
  while ((s + 1) * (s + 2) / 2 <= i) ++s;  // s = l - lx
  int lz = i - s * (s + 1) / 2;
  int lx = l - s;
  int ly = s - lz;
  return lx | (ly << 8) | (lz << 16);
}
// inverse of hidx: packed t | u<<8 | v<<16
__host__ __device__ constexpr int herm_of(int idx) {
  int n = 0;
  while (nherm(n) <= idx) ++n;
  int i = idx - (n > 0 ? nherm(n - 1) : 0);
  return cart_of(n, i);
}
__host__ __device__ constexpr int px(int p) { return p & 0xff; }
__host__ __device__ constexpr int py(int p) { return (p >> 8) & 0xff; }
__host__ __device__ constexpr int pz(int p) { return (p >> 16) & 0xff; }

constexpr int MAXL = 5;                 // highest shell angular momentum (AUTO2E_MAXL)
constexpr int BOYS_NINT = 500;          // boys.f90:21,45 (BOYS_ACCURACY=3)
constexpr double BOYS_XASYM = 25.0;     // boys.f90:20,44
constexpr int BOYS_MAXM = 4 * MAXL;     // boys_init(4*AUTO2E_MAXL)

// Spherical transform coefficients U(l) for l = 2..MAXL, rows m=-l..l,
// row-major (2l+1) x ncart(l), packed consecutively from l = 2.
constexpr int spher_off(int l) {
  int o = 0;
  for (int k = 2; k < l; ++k) o += (2 * k + 1) * ncart(k);
  return o;
}
constexpr int SPHER_TOTAL = spher_off(MAXL + 1);

// ---------------------------------------------------------------- basis on device
struct DevShell {
  double x, y, z;
  int l;       // angular momentum
  int np;      // primitives
  int prim0;   // offset into exps/cntr
  int nang;    // spherical (l>=2) or Cartesian count
  int norm0;   // offset into norms
  int pad;
};

// host helpers (gi_tables.cpp)
void boys_table_host(int maxm, std::vector<double> &table);   // (5, 500, 0:maxm)
void spher_u_host(int l, double *U);                            // (2l+1) x ncart(l)

}  // namespace gi
