// Host-side tables for the device ERI kernels.
//
// Boys table: 5-point Chebyshev coefficients of F_m on 500 intervals of
// [0, 25) for m = 0..maxm, built from an accurate evaluation of F_m at the
// Chebyshev nodes of every interval (the table layout, nodes and coefficient
// formula of boys.f90:57-86 and cheb_nodes/cheb_coeffs chebinterp.f90:9-56;
// F_m at the nodes by the series below).  The device
// interpolates only at the top order a block needs and recurses downwards,
// exactly as the reference's fm() does, so ERIs agree with the reference's
// tabulation error included.
//
// Solid harmonics: real, unit-sphere normalized, m = -l..l (Helgaker
// 6.4.47-50; AutoSpherTransf.py:62-158).
#include <cmath>
#include <cstdlib>
#include "gi_common.h"

namespace gi {

static const double PI_ = 3.141592653589793;

// F_m(x) at the table's Chebyshev nodes, x in [0, 25): the series
//   F_m(x) = exp(-x) sum_{i>=0} (2x)^i / ((2m+1)(2m+3)...(2m+2i+1))
// (every term positive, so no cancellation), summed in extended precision
// until a term falls below 2^-70 of the sum.  Accurate to the last bit of a
// double over the whole table range.
static double boys_accurate(long m, double x) {
  const long double twox = 2.0L * (long double)x;
  long double term = 1.0L / (long double)(2 * m + 1), sum = term;
  for (long i = 1; i < 100000 && term > sum * 0x1p-70L; ++i) {
    term *= twox / (long double)(2 * m + 2 * i + 1);
    sum += term;
  }
  return (double)(std::exp(-(long double)x) * sum);
}

void boys_table_host(int maxm, std::vector<double> &table) {
  const int NPT = 5;
  table.assign((size_t)NPT * BOYS_NINT * (maxm + 1), 0.0);
  const double dx = BOYS_XASYM / (double)BOYS_NINT;
  for (int k = 0; k <= maxm; ++k)
    for (int i = 1; i <= BOYS_NINT; ++i) {
      double a = (double)(i - 1) * dx, b = (double)i * dx;
      double w0 = PI_ / (double)NPT, w1 = (b - a) / 2.0, w2 = (b + a) / 2.0;
      double fs[NPT];
      for (int j = 1; j <= NPT; ++j) {
        double t = w0 * ((double)j - 0.5);
        fs[j - 1] = boys_accurate(k, w1 * std::cos(t) + w2);
      }
      double *c = table.data() + NPT * ((i - 1) + (size_t)BOYS_NINT * k);
      for (int j = 1; j <= NPT; ++j) {
        double s = 0.0, wj = (double)(j - 1);
        for (int q = 1; q <= NPT; ++q) s = s + fs[q - 1] * std::cos(w0 * wj * ((double)q - 0.5));
        c[j - 1] = 2.0 / (double)NPT * s;
      }
    }
}

static double fct(int n) {
  double r = 1.0;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}
static double bnm(int n, int k) {
  if (k < 0 || n < 0 || k > n) return 0.0;
  return fct(n) / (fct(k) * fct(n - k));
}

void spher_u_host(int l, double *U) {
  const int nc = ncart(l);
  for (int i = 0; i < (2 * l + 1) * nc; ++i) U[i] = 0.0;
  for (int m = -l; m <= l; ++m) {
    int am = std::abs(m), vm = m >= 0 ? 0 : 1;
    double pref = std::sqrt((2.0 * l + 1.0) / (4.0 * PI_)) * std::sqrt(fct(l + am)) *
                  std::sqrt(fct(l - am)) * (m == 0 ? 1.0 : std::sqrt(2.0)) /
                  (std::pow(2.0, am) * fct(l));
    for (int t = 0; t <= (l - am) / 2; ++t)
      for (int u = 0; u <= t; ++u)
        for (int k = 0; k <= (am - vm) / 2; ++k) {
          int v = 2 * k + vm;
          int lx = 2 * t + am - 2 * u - v, ly = 2 * u + v, lz = l - 2 * t - am;
          double num = bnm(l, t) * bnm(l - t, am + t) * bnm(t, u) * bnm(am, v);
          double den = (((t + (v - vm) / 2) % 2) ? -1.0 : 1.0) * std::pow(4.0, t);
          U[(m + l) * nc + cidx(lx, ly, lz)] += pref * (num / den);
        }
  }
}

}  // namespace gi
