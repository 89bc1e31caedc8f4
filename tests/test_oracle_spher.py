"""Pins of the solid-harmonic transform U(l) (oracle or_spher_u, the same
table the library folds into its Hermite tables, csrc/gi_tables.cpp):

* equal to the coefficients the reference's generated code states,
  auto1e_SpherTransf_U_<l> (auto2e_SpherTransf.f90:3938-4130), for l = 2..5
  (tests/golden/spher_u_reference.json, written by make_spher_u.py from the
  text of that file);
* equal to an independent derivation in exact arithmetic: real solid
  harmonics r^l Y_lm normalized on the unit sphere, m = -l..l, no
  Condon-Shortley phase (m > 0 ~ Re (x + iy)^m, m < 0 ~ Im (x + iy)^|m|),
  expanded in the Cartesian monomials x^lx y^ly z^lz in the Auto2e order."""
import json
import os

import numpy as np
import pytest
import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))


def cart_order(l):
    return [(lx, l - lx - lz, lz) for lx in range(l, -1, -1) for lz in range(0, l - lx + 1)]


def solid_harmonics(l):
    x, y, z, t = sp.symbols("x y z t", real=True)
    r2 = x ** 2 + y ** 2 + z ** 2
    Pl = sp.legendre(l, t)
    rows = []
    for m in range(-l, l + 1):
        am = abs(m)
        dP = sp.Poly(sp.diff(Pl, t, am), t)   # d^m P_l / dt^m, degree d = l - m, parity of d
        d = l - am
        # r^d dP(z / r) = sum_k c_k z^k (r^2)^((d - k) / 2), d - k even
        poly = sp.expand(sum(c * z ** k * r2 ** ((d - k) // 2) for (k,), c in dP.terms()))
        w = sp.expand((x + sp.I * y) ** am)
        ang = sp.re(w) if m >= 0 else sp.im(w)
        norm = sp.sqrt(sp.Rational(2 * l + 1, 4) / sp.pi * sp.factorial(l - am) / sp.factorial(l + am))
        if m != 0:
            norm *= sp.sqrt(2)
        S = sp.Poly(sp.expand(norm * poly * ang), x, y, z)
        rows.append([float(S.coeff_monomial(x ** a * y ** b * z ** c)) for a, b, c in cart_order(l)])
    return np.array(rows)


@pytest.mark.parametrize("l", [2, 3, 4, 5])
def test_u_equals_reference_literals(oracle_mod, l):
    ref = np.array(json.load(open(os.path.join(HERE, "golden", "spher_u_reference.json")))["U"][str(l)])
    U = oracle_mod.spher_u(l)
    assert U.shape == ref.shape == (2 * l + 1, (l + 1) * (l + 2) // 2)
    assert np.abs(U - ref).max() <= 1e-15 * np.abs(ref).max()


@pytest.mark.parametrize("l", [0, 1, 2, 3, 4, 5])
def test_u_equals_exact_solid_harmonics(oracle_mod, l):
    S = solid_harmonics(l)
    if l < 2:
        # s and p are not transformed (p = x, y, z); the normalized solid
        # harmonics are the same functions up to a constant and order
        assert S.shape[0] == 2 * l + 1
        return
    U = oracle_mod.spher_u(l)
    assert np.abs(U - S).max() <= 2e-15 * np.abs(S).max()


def test_planted_sign_is_detected():
    """The exact derivation distinguishes the sign conventions: with the
    Condon-Shortley phase the m = +-1 rows of d change sign."""
    S = solid_harmonics(2)
    ref = np.array(json.load(open(os.path.join(HERE, "golden", "spher_u_reference.json")))["U"]["2"])
    flipped = S.copy()
    flipped[[1, 3]] *= -1
    assert np.abs(S - ref).max() < 1e-15 < np.abs(flipped - ref).max()
