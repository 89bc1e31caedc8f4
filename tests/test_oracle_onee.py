"""The oracle's one-electron integrals (S, T, V_ne of ints1e_gammcor_H0),
pinned by Gauss-Hermite quadrature (exact for these integrands), closed forms
for s functions, and the oracle's ERIs with a tight s function standing in
for a point charge."""
import math

import numpy as np
import pytest

import pyoracle
from gigpu.basis import build_basis, cart_components, spher_u


@pytest.fixture(scope="module", autouse=True)
def _built():
    pyoracle.build()


def two_shell_basis(la, lb, a, b, spher=True):
    coords = np.array([[0.1, -0.3, 0.2], [0.9, 0.5, -0.6]])
    return build_basis(["A", "B"], coords, {"A": [(la, [a], [1.0])], "B": [(lb, [b], [1.0])]}, spher_ao=spher)


def gh_1d(n=24):
    x, w = np.polynomial.hermite.hermgauss(n)
    return x, w


def prim_cart(la, lb, a, b, A, B):
    """S and T between unnormalized Cartesian primitives by Gauss-Hermite
    quadrature, one dimension at a time."""
    p = a + b
    P = (a * A + b * B) / p
    K = math.exp(-a * b / p * np.dot(A - B, A - B))
    s, w = gh_1d()
    ca, cb = cart_components(la), cart_components(lb)
    S = np.zeros((len(ca), len(cb)))
    T = np.zeros((len(ca), len(cb)))
    for i, (ax, ay, az) in enumerate(ca):
        for j, (bx, by, bz) in enumerate(cb):
            ov, kin = [], []
            for d, (ea, eb) in enumerate(((ax, bx), (ay, by), (az, bz))):
                x = P[d] + s / math.sqrt(p)
                fa = (x - A[d]) ** ea
                xb = x - B[d]
                fb = xb ** eb
                # d^2/dx^2 of xb^eb exp(-b xb^2), divided by the Gaussian
                d2 = (eb * (eb - 1) * xb ** (eb - 2) if eb >= 2 else 0.0) - 2 * b * (2 * eb + 1) * xb ** eb \
                    + 4 * b * b * xb ** (eb + 2)
                ov.append(np.sum(w * fa * fb) / math.sqrt(p))
                kin.append(np.sum(w * fa * d2) / math.sqrt(p))
            S[i, j] = K * ov[0] * ov[1] * ov[2]
            T[i, j] = -0.5 * K * (kin[0] * ov[1] * ov[2] + ov[0] * kin[1] * ov[2] + ov[0] * ov[1] * kin[2])
    return S, T


def to_ao(X, basis, la, lb):
    """Cartesian primitive block -> AO block: contraction coefficient, U for
    spherical l >= 2, AO normalization factors."""
    def side(l, p):
        c = basis.cntr_coeffs[p, 0]
        n = basis.norm_factors[p]
        if basis.spher_ao and l >= 2:
            U = spher_u(l)
            return c * n[:U.shape[0], None] * U
        k = len(cart_components(l))
        return c * np.diag(n[:k])
    Ua = side(la, 0)
    Ub = side(lb, 1)
    return Ua @ X @ Ub.T


@pytest.mark.parametrize("la,lb", [(0, 0), (1, 0), (1, 1), (2, 1), (2, 2), (3, 2), (3, 3)])
@pytest.mark.parametrize("spher", [True, False])
def test_overlap_kinetic_against_quadrature(la, lb, spher):
    a, b = 0.8, 1.3
    basis = two_shell_basis(la, lb, a, b, spher)
    o = pyoracle.Oracle(basis)
    S, T, _ = o.one_electron(np.zeros(2))
    A, B = basis.atom_coords
    Sc, Tc = prim_cart(la, lb, a, b, A, B)
    na = basis.shell_nang(0)
    Sab = to_ao(Sc, basis, la, lb)
    Tab = to_ao(Tc, basis, la, lb)
    assert np.abs(S[:na, na:] - Sab).max() <= 1e-13 * max(1.0, np.abs(Sab).max())
    assert np.abs(T[:na, na:] - Tab).max() <= 1e-13 * max(1.0, np.abs(Tab).max())
    assert np.allclose(S, S.T, atol=0) and np.allclose(T, T.T, atol=1e-15)
    if spher:   # normalized AOs
        assert np.abs(np.diag(S) - 1.0).max() <= 1e-13


def test_s_function_closed_forms():
    a = 0.9
    basis = build_basis(["H", "H"], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.7]]), {"H": [(0, [a], [1.0])]})
    o = pyoracle.Oracle(basis)
    S, T, V = o.one_electron(np.array([1.0, 0.0]))
    R = 1.7
    assert abs(S[0, 1] - math.exp(-a / 2 * R * R)) <= 1e-14
    assert abs(T[0, 0] - 1.5 * a) <= 1e-13
    mu = a / 2
    assert abs(T[0, 1] - mu * (3 - 2 * mu * R * R) * S[0, 1]) <= 1e-13
    # phi^2 is a unit Gaussian charge of exponent 2a: potential 2 sqrt(2a/pi) at its centre,
    # erf(sqrt(2a) R)/R at distance R
    assert abs(V[0, 0] + 2 * math.sqrt(2 * a / math.pi)) <= 1e-12
    assert abs(V[1, 1] + math.erf(math.sqrt(2 * a) * R) / R) <= 1e-12


@pytest.mark.parametrize("la,lb", [(1, 0), (1, 1), (2, 1), (2, 2), (3, 1), (3, 3)])
def test_nuclear_attraction_against_tight_charge_eris(la, lb):
    """V_ab for a unit charge at C equals -(ab|xx) with x an s function of
    exponent gamma -> infinity at C (x^2 is a unit charge of width
    1/sqrt(2 gamma))."""
    gamma = 1e9
    C = np.array([0.4, -0.1, 0.8])
    coords = np.array([[0.1, -0.3, 0.2], [0.9, 0.5, -0.6], C])
    shells = {"A": [(la, [0.8], [1.0])], "B": [(lb, [1.3], [1.0])], "X": [(0, [gamma], [1.0])]}
    basis = build_basis(["A", "B", "X"], coords, shells)
    o = pyoracle.Oracle(basis)
    _, _, V = o.one_electron(np.array([0.0, 0.0, 1.0]))
    na, nb = basis.shell_nang(0), basis.shell_nang(1)
    ref = -o.eri(0, 1, 2, 2)[:, :, 0, 0]
    assert np.abs(V[:na, na:na + nb] - ref).max() <= 1e-7 * np.abs(ref).max()
