"""Pins of the oracle's Singleton quicksort (DSORT0, common/sort.f90:19-330).

The permutation of tied keys fixes the Cholesky pivot order, so the oracle's
C version is compared, permutation included, against an independent Python
transcription that follows the Fortran statement labels (110..190) one by
one as a state machine."""
import numpy as np
import pytest


def dsort0_labels(dx, dy, kflag):
    dx = [float(v) for v in dx]
    dy = list(dy)
    n = len(dx)
    if n <= 1:
        return dx, dy
    if kflag <= -1:
        dx = [-v for v in dx]
    X = [None] + dx
    Y = [None] + dy
    il = [0] * 22
    iu = [0] * 22
    m, i, j, r = 1, 1, n, 0.375
    k = l = ij = 0
    t = ty = None
    label = 110
    while True:
        if label == 110:
            if i == j:
                label = 150
                continue
            r = r + 3.90625e-2 if r <= 0.5898437 else r - 0.21875
            label = 120
        elif label == 120:
            k = i
            ij = i + int((j - i) * r)
            t, ty = X[ij], Y[ij]
            if X[i] > t:
                X[ij], X[i], t = X[i], t, X[i]
                Y[ij], Y[i], ty = Y[i], ty, Y[i]
            l = j
            if X[j] < t:
                X[ij], X[j], t = X[j], t, X[j]
                Y[ij], Y[j], ty = Y[j], ty, Y[j]
                if X[i] > t:
                    X[ij], X[i], t = X[i], t, X[i]
                    Y[ij], Y[i], ty = Y[i], ty, Y[i]
            label = 130
        elif label == 130:
            l -= 1
            if X[l] > t:
                continue
            label = 140
        elif label == 140:
            k += 1
            if X[k] < t:
                continue
            if k <= l:
                X[l], X[k] = X[k], X[l]
                Y[l], Y[k] = Y[k], Y[l]
                label = 130
                continue
            if l - i > j - k:
                il[m], iu[m], i = i, l, k
            else:
                il[m], iu[m], j = k, j, l
            m += 1
            label = 160
        elif label == 150:
            m -= 1
            if m == 0:
                label = 190
                continue
            i, j = il[m], iu[m]
            label = 160
        elif label == 160:
            if j - i >= 1:
                label = 120
                continue
            if i == 1:
                label = 110
                continue
            i -= 1
            label = 170
        elif label == 170:
            i += 1
            if i == j:
                label = 150
                continue
            t, ty = X[i + 1], Y[i + 1]
            if X[i] <= t:
                continue
            k = i
            label = 180
        elif label == 180:
            X[k + 1], Y[k + 1] = X[k], Y[k]
            k -= 1
            if t < X[k]:
                continue
            X[k + 1], Y[k + 1] = t, ty
            label = 170
        elif label == 190:
            out = X[1:]
            if kflag <= -1:
                out = [-v for v in out]
            return out, Y[1:]


@pytest.mark.parametrize("kflag", [2, -2])
def test_dsort0_matches_label_transcription(oracle_mod, kflag):
    rng = np.random.default_rng(1)
    for trial in range(300):
        n = int(rng.integers(1, 400))
        if trial % 3 == 0:
            x = rng.integers(0, 5, n).astype(float)          # many ties
        elif trial % 3 == 1:
            x = np.sort(rng.random(n))[::-1] if kflag < 0 else np.sort(rng.random(n))   # presorted
            x[rng.integers(0, n, max(1, n // 7))] = 0.5      # with ties
        else:
            x = rng.normal(size=n)
        y = np.arange(1, n + 1)
        xs, ys = oracle_mod.dsort0(x, y, kflag)
        xr, yr = dsort0_labels(x, y, kflag)
        assert np.array_equal(xs, np.array(xr)), trial
        assert np.array_equal(ys, np.array(yr)), trial
        assert np.all(np.diff(xs) >= 0) if kflag > 0 else np.all(np.diff(xs) <= 0)
        assert np.array_equal(np.sort(ys), y)
        assert np.array_equal(x[ys - 1], xs)
