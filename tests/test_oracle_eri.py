"""Pins of the oracle ERIs (restating the Auto2e McMurchie-Davidson + HRR
engine) against mathematics that does not share its algebra."""
import math

import numpy as np
import pytest

from gigpu import systems
from gigpu.basis import build_basis, spher_u, cart_components
import independent_eri as ind


@pytest.fixture(scope="module")
def water(oracle_mod):
    b = systems.water_system(2, 6.0, "large", seed=11)
    return b, oracle_mod.Oracle(b)


def _shells_of(b, l, k=0):
    s = [i for i in range(b.nshells) if b.shell_l(i) == l]
    return s[k % len(s)]


def test_ssss_closed_form(water):
    b, o = water
    ss = [i for i in range(b.nshells) if b.shell_l(i) == 0][:6]
    rng = np.random.default_rng(3)
    for _ in range(20):
        q = rng.choice(ss, 4)
        V = o.eri(*q)[0, 0, 0, 0]
        ref = ind.shell_quartet(b, q)[0, 0, 0, 0]
        assert abs(V - ref) <= 5e-12 * max(1.0, abs(ref))


def test_against_obara_saika(water):
    """Random quartets up to (ff|ff) against the independent OS/HGP
    evaluator with exact Boys values."""
    b, o = water
    rng = np.random.default_rng(5)
    by_l = {l: [i for i in range(b.nshells) if b.shell_l(i) == l] for l in range(4)}
    combos = [(0, 0, 0, 0), (1, 0, 1, 0), (1, 1, 1, 1), (2, 0, 0, 0), (2, 1, 2, 0), (2, 2, 1, 1),
              (2, 2, 2, 2), (3, 0, 1, 0), (0, 3, 2, 1), (3, 2, 1, 3), (3, 3, 3, 3), (1, 2, 3, 0)]
    for ls in combos:
        q = [int(rng.choice(by_l[l])) for l in ls]
        V = o.eri(*q)
        R = ind.shell_quartet(b, q)
        err = np.abs(V - R).max()
        assert err <= 2e-11 * max(1.0, np.abs(R).max()), (ls, q, err)


@pytest.mark.parametrize("spher", [True, False], ids=["spherical", "cartesian"])
def test_g_and_h_shells_against_obara_saika(oracle_mod, spher):
    """Classes with g and h shells (AUTO2E_MAXL = 5) against the independent
    evaluator, spherical and Cartesian."""
    b = systems.water_system(1, 6.0, "gh", seed=3, spher_ao=spher)
    o = oracle_mod.Oracle(b)
    rng = np.random.default_rng(7)
    by_l = {l: [i for i in range(b.nshells) if b.shell_l(i) == l] for l in range(6)}
    combos = [(4, 0, 0, 0), (4, 1, 2, 0), (5, 0, 1, 0), (4, 4, 0, 0), (5, 4, 2, 1), (4, 3, 4, 2), (5, 5, 1, 0),
              (0, 5, 3, 3), (5, 2, 5, 0)]
    for ls in combos:
        q = [int(rng.choice(by_l[l])) for l in ls]
        V = o.eri(*q)
        R = ind.shell_quartet(b, q)
        err = np.abs(V - R).max()
        assert err <= 2e-11 * max(1.0, np.abs(R).max()), (ls, q, err)


@pytest.mark.parametrize("spher", [True, False], ids=["spherical", "cartesian"])
def test_contracted_shells_against_obara_saika(oracle_mod, spher):
    """Segmented contractions (3- and 2-primitive shells mixed with single
    primitives) in spherical and Cartesian AOs against the independent
    evaluator: a contraction coefficient applied twice or not at all, a
    primitive loop bound off by one, or a Cartesian norm factor on the wrong
    component fails it."""
    b = systems.water_system(1, 6.0, "large", seed=31, contracted=True, spher_ao=spher)
    o = oracle_mod.Oracle(b)
    multi = {l: [i for i in range(b.nshells) if b.shell_l(i) == l and b.nprimitives[b.shell_params_idx[i] - 1] > 1]
             for l in range(4)}
    single = {l: [i for i in range(b.nshells) if b.shell_l(i) == l] for l in range(4)}
    rng = np.random.default_rng(6)
    combos = [(0, 0, 0, 0), (1, 0, 1, 1), (2, 1, 0, 2), (3, 0, 1, 0), (2, 2, 3, 1), (1, 3, 2, 0)]
    for k, ls in enumerate(combos):
        q = [int(rng.choice(multi[l] if (j + k) % 2 == 0 and multi[l] else single[l])) for j, l in enumerate(ls)]
        V = o.eri(*q)
        R = ind.shell_quartet(b, q)
        assert V.shape == R.shape
        err = np.abs(V - R).max()
        assert err <= 2e-11 * max(1.0, np.abs(R).max()), (ls, q, err)


def test_kappa_attenuated_ssss(oracle_mod):
    # erf(omega r12)/r12: alpha -> PQ/(P+Q+kappa PQ), closed form for s functions
    b = build_basis(["H", "H"], np.array([[0.0, 0.0, 0.0], [0.0, 0.3, 1.4]]),
                    {"H": [(0, [0.7], [1.0]), (0, [3.1], [1.0])]})
    o = oracle_mod.Oracle(b)
    kappa = 1.0 / 0.4 ** 2
    V = o.eri(0, 3, 1, 2, kappa=kappa)[0, 0, 0, 0]
    def g(s):
        p = b.shell_params_idx[s] - 1
        return b.atom_coords[b.shell_centers[s] - 1], b.exponents[p, 0], b.cntr_coeffs[p, 0] * b.norm_spher[p, 0]
    (A, a, ca), (B, bb, cb), (C, c, cc), (D, d, cd) = g(0), g(3), g(1), g(2)
    P, Q = a + bb, c + d
    Rp, Rq = (a * A + bb * B) / P, (c * C + d * D) / Q
    K = math.exp(-a * bb / P * np.sum((A - B) ** 2) - c * d / Q * np.sum((C - D) ** 2))
    den = P + Q + kappa * P * Q
    al = P * Q / den
    x = al * np.sum((Rp - Rq) ** 2)
    F0 = 0.5 * math.sqrt(math.pi / x) * math.erf(math.sqrt(x))
    ref = 2 * math.pi ** 2.5 / (P * Q * math.sqrt(den)) * K * F0 * ca * cb * cc * cd
    assert abs(V - ref) < 1e-12


def test_permutational_symmetry(water):
    b, o = water
    q = [_shells_of(b, 2, 1), _shells_of(b, 1, 3), _shells_of(b, 3), _shells_of(b, 0, 5)]
    V = o.eri(*q)
    assert np.allclose(o.eri(q[1], q[0], q[2], q[3]), V.transpose(1, 0, 2, 3), rtol=0, atol=1e-14)
    assert np.allclose(o.eri(q[0], q[1], q[3], q[2]), V.transpose(0, 1, 3, 2), rtol=0, atol=1e-14)
    assert np.allclose(o.eri(q[2], q[3], q[0], q[1]), V.transpose(2, 3, 0, 1), rtol=0, atol=1e-14)


def _rotation(seed):
    rng = np.random.default_rng(seed)
    qq = rng.normal(size=4)
    w, x, y, z = qq / np.linalg.norm(qq)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_translation_and_rotation_invariance(oracle_mod):
    """Block Frobenius norms of complete spherical shells are invariant under
    rigid motions (real solid harmonics of one l rotate orthogonally); a
    wrong coefficient or sign in a d/f transform breaks this."""
    shells = {"O": [(0, [1.3], [1.0]), (1, [0.8], [1.0]), (2, [0.9], [1.0]), (3, [0.7], [1.0])],
              "H": [(1, [0.5], [1.0]), (2, [1.1], [1.0])]}
    xyz = np.array([[0.1, 0.2, -0.3], [1.2, -0.4, 0.9], [-0.8, 1.1, 0.5]])
    b0 = build_basis(["O", "H", "H"], xyz, shells)
    o0 = oracle_mod.Oracle(b0)
    Rm = _rotation(7)
    b1 = build_basis(["O", "H", "H"], xyz @ Rm.T + np.array([3.0, -1.0, 2.0]), shells)
    o1 = oracle_mod.Oracle(b1)
    rng = np.random.default_rng(2)
    for _ in range(15):
        q = rng.integers(0, b0.nshells, 4)
        n0 = np.linalg.norm(o0.eri(*q))
        n1 = np.linalg.norm(o1.eri(*q))
        assert abs(n0 - n1) <= 1e-11 * max(1.0, n0)


def test_unit_normalization_via_self_overlap():
    """NormFactorsSpher/Cart make every AO unit-normalized: check with an
    independent Gauss-Hermite quadrature of the squared function."""
    from gigpu.basis import normalize_shell
    xs, ws = np.polynomial.hermite.hermgauss(40)
    for l in range(4):
        alpha = 0.9
        cntr, ncart_, nsph = normalize_shell(l, np.array([alpha]), np.array([1.0]))
        # integral of x^(2i) exp(-2 alpha x^2): substitute t = sqrt(2 alpha) x
        def mom(i):
            s = np.sqrt(2 * alpha)
            return float(np.sum(ws * (xs / s) ** (2 * i)) / s)
        comps = cart_components(l)
        S = np.zeros((len(comps), len(comps)))
        for i, a in enumerate(comps):
            for j, c in enumerate(comps):
                e = [a[k] + c[k] for k in range(3)]
                if any(v % 2 for v in e):
                    continue
                S[i, j] = cntr[0] ** 2 * mom(e[0] // 2) * mom(e[1] // 2) * mom(e[2] // 2)
        for i in range(len(comps)):
            assert abs(S[i, i] * ncart_[i] ** 2 - 1.0) < 1e-12
        U = spher_u(l) if l >= 2 else np.eye(len(comps))
        T = U @ S @ U.T
        for m in range(T.shape[0]):
            assert abs(T[m, m] * nsph[m] ** 2 - 1.0) < 1e-12
        # distinct m are orthogonal (real solid harmonics)
        off = T * np.outer(nsph, nsph)
        assert np.abs(off - np.diag(np.diag(off))).max() < 1e-12


def test_spher_u_oracle_matches_host(oracle_mod):
    for l in range(2, 6):
        assert np.abs(oracle_mod.spher_u(l) - spher_u(l)).max() < 1e-15


def test_d_harmonics_convention():
    """m = -2..2 for d: xy, yz, 3z^2-r^2, xz, x^2-y^2 (auto2e_SpherTransf.f90:3938-3946)."""
    U = spher_u(2)
    # Cartesian order xx, xy, xz, yy, yz, zz
    assert np.allclose(U[0], [0, 1.0925484305920792, 0, 0, 0, 0])
    assert np.allclose(U[1], [0, 0, 0, 0, 1.0925484305920792, 0])
    assert np.allclose(U[2], [-0.3153915652525201, 0, 0, -0.3153915652525201, 0, 0.6307831305050402])
    assert np.allclose(U[3], [0, 0, 1.0925484305920792, 0, 0, 0])
    assert np.allclose(U[4], [0.54627421529603959, 0, 0, -0.54627421529603959, 0, 0])


def test_schwarz_bound(water):
    b, o = water
    rng = np.random.default_rng(9)
    for _ in range(10):
        q = rng.integers(0, b.nshells, 4)
        V = o.eri(*q)
        Dab = np.einsum("abab->ab", o.eri(q[0], q[1], q[0], q[1]))
        Dcd = np.einsum("cdcd->cd", o.eri(q[2], q[3], q[2], q[3]))
        bound = np.sqrt(np.abs(Dab))[:, :, None, None] * np.sqrt(np.abs(Dcd))[None, None, :, :]
        assert np.all(np.abs(V) <= bound * (1 + 1e-10) + 1e-14)
