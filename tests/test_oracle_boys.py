"""Pins of the oracle's Boys function (restating boys.f90) against
mathematics: mpmath evaluation of the defining integral, the erf closed form
of F_0, the truncated asymptotic branch and the recursion identities."""
import math

import mpmath
import numpy as np
import pytest

mpmath.mp.dps = 40


def exact(m, x):
    if x == 0:
        return 1.0 / (2 * m + 1)
    return float(mpmath.hyp1f1(m + 0.5, m + 1.5, -x) / (2 * m + 1))


def test_fm_reference_relative_error(oracle_mod):
    # boys.f90:96-97 states a relative error below 4e-15 for all x >= 0
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(0, 40, 200), [0.0, 1e-12, 0.999, 1.0, 24.999, 25.0, 34.9, 35.0, 50.0]])
    worst = 0.0
    for x in xs:
        for m in (0, 1, 3, 7, 12, 20):
            e = exact(m, x)
            worst = max(worst, abs(oracle_mod.boys_fm_reference(m, x) - e) / e)
    assert worst < 4e-15


def test_table_interpolation_error(oracle_mod):
    # tabulated F_m with downward recursion: max unsigned error 4.306e-12 (boys.f90:30)
    rng = np.random.default_rng(1)
    worst = 0.0
    for x in rng.uniform(0, 25, 150):
        for top in (0, 4, 8, 12):
            F = oracle_mod.boys_fm(top, x)
            for m in range(top + 1):
                worst = max(worst, abs(F[m] - exact(m, x)))
    assert worst < 4.5e-12


def test_f0_closed_form(oracle_mod):
    for x in (1e-3, 0.5, 2.0, 10.0, 24.0):
        e = 0.5 * math.sqrt(math.pi / x) * math.erf(math.sqrt(x))
        assert abs(oracle_mod.boys_fm(0, x)[0] - e) < 2e-12


def test_asymptotic_branch_is_truncated(oracle_mod):
    # x >= 25: F_0 = 1/2 sqrt(pi/x) without erf, upward recursion (boys.f90:263-280)
    for x in (25.0, 31.7, 80.0):
        F = oracle_mod.boys_fm(6, x)
        f = 0.5 * math.sqrt(math.pi / x)
        assert F[0] == f
        beta, a, ex = 1.0 / (2.0 * x), 1.0, math.exp(-x)
        for m in range(1, 7):
            f = beta * (a * f - ex)
            a += 2.0
            assert F[m] == f


def test_downward_recursion_identity(oracle_mod):
    # F_{m-1} = (2x F_m + e^-x)/(2m-1) holds exactly for the returned array
    for x in (0.3, 7.7, 19.2):
        F = oracle_mod.boys_fm(10, x)
        for m in range(10, 0, -1):
            assert abs(F[m - 1] - (2 * x * F[m] + math.exp(-x)) / (2 * m - 1)) <= 1e-15 * abs(F[m - 1])


def test_zero_argument(oracle_mod):
    # at x = 0 the downward recursion is exact (2x F_m vanishes); only the
    # interpolated top order carries the table error
    F = oracle_mod.boys_fm(8, 0.0)
    for m in range(8):
        assert abs(F[m] - 1.0 / (2 * m + 1)) < 1e-16
    assert abs(F[8] - 1.0 / 17) < 4.3e-12
