// Singleton's quicksort (CACM Algorithm 347) as the reference's DSORT0
// (common/sort.f90:19-330) applies it.  The Cholesky pivot order depends on
// the exact permutation of tied keys, so the steps are the published ones:
// median-of-three on a moving fraction r of the segment, a 21-deep segment
// stack, straight insertion for segments past the first (the element left of
// such a segment is a sentinel).  kflag = +-2 carries dy; negative = descending.
namespace gi {

void dsort0(double *dx, long long *dy, long long n, int kflag) {
  if (n <= 1) return;
  if (kflag <= -1)
    for (long long i = 0; i < n; ++i) dx[i] = -dx[i];
  double *X = dx - 1;
  long long *Y = dy - 1;
  long long il[22], iu[22];
  long long m = 1, i = 1, j = n, k, l, ij, ty = 0;
  double r = 0.375, t;
  enum { NEW_FRACTION, PARTITION, POP, CHECK, INSERTION, DONE } st = NEW_FRACTION;
  for (;;) {
    switch (st) {
      case NEW_FRACTION:
        if (i == j) { st = POP; break; }
        r = (r <= 0.5898437) ? r + 3.90625e-2 : r - 0.21875;
        st = PARTITION;
        break;
      case PARTITION: {
        k = i;
        ij = i + (long long)((double)(j - i) * r);
        t = X[ij];
        ty = Y[ij];
        if (X[i] > t) { X[ij] = X[i]; X[i] = t; t = X[ij]; Y[ij] = Y[i]; Y[i] = ty; ty = Y[ij]; }
        l = j;
        if (X[j] < t) {
          X[ij] = X[j]; X[j] = t; t = X[ij]; Y[ij] = Y[j]; Y[j] = ty; ty = Y[ij];
          if (X[i] > t) { X[ij] = X[i]; X[i] = t; t = X[ij]; Y[ij] = Y[i]; Y[i] = ty; ty = Y[ij]; }
        }
        for (;;) {
          do { --l; } while (X[l] > t);
          do { ++k; } while (X[k] < t);
          if (k > l) break;
          double tt = X[l]; X[l] = X[k]; X[k] = tt;
          long long tty = Y[l]; Y[l] = Y[k]; Y[k] = tty;
        }
        if (l - i > j - k) { il[m] = i; iu[m] = l; i = k; }
        else { il[m] = k; iu[m] = j; j = l; }
        ++m;
        st = CHECK;
        break;
      }
      case POP:
        --m;
        if (m == 0) { st = DONE; break; }
        i = il[m];
        j = iu[m];
        st = CHECK;
        break;
      case CHECK:
        if (j - i >= 1) { st = PARTITION; break; }
        if (i == 1) { st = NEW_FRACTION; break; }
        --i;
        st = INSERTION;
        break;
      case INSERTION:
        for (;;) {
          ++i;
          if (i == j) break;
          t = X[i + 1];
          ty = Y[i + 1];
          if (X[i] <= t) continue;
          k = i;
          do { X[k + 1] = X[k]; Y[k + 1] = Y[k]; --k; } while (t < X[k]);
          X[k + 1] = t;
          Y[k + 1] = ty;
        }
        st = POP;
        break;
      case DONE:
        if (kflag <= -1)
          for (long long q = 0; q < n; ++q) dx[q] = -dx[q];
        return;
    }
  }
}

}  // namespace gi
