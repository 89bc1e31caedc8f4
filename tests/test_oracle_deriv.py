"""Pins of the derivative-ERI oracle (or_eri_deriv_a) and of the gradient
it gives (or_coulomb_grad), which check the library's nuclear gradient.

* d(ab|cd)/dA against central differences (h = 1e-4 bohr) of the oracle's
  own ERIs with shell a alone moved (a copy of shell a on a private atom
  in the first position), for s through g on shell a, contracted and uncontracted,
  spherical and Cartesian, Kappa = 0 and 0.4;
* translational invariance: the derivatives with respect to the four
  centres of a quartet sum to zero;
* dE_J/dR from derivative ERIs against central differences of the oracle's
  Coulomb energy, and its rows sum to zero.
Both terms of the derivative (the raised shell and the lowered one) are of
the size of the result, so dropping either, or a wrong sign or factor,
fails the difference checks."""
import dataclasses

import numpy as np
import pytest

from gigpu import systems
from gigpu.basis import build_basis
from grad_helpers import fd_gradient, random_density, tiny_basis


def _private_copy(basis, s):
    """Copy of basis with one more shell: a duplicate of shell s on a new
    atom at the same position (so that only the first position of a quartet
    moves even when s appears at other positions too)."""
    at = basis.shell_centers[s] - 1
    coords = np.vstack([basis.atom_coords, basis.atom_coords[at:at + 1]])
    centers = np.append(basis.shell_centers, coords.shape[0]).astype(np.int64)
    params = np.append(basis.shell_params_idx, basis.shell_params_idx[s]).astype(np.int64)
    return dataclasses.replace(basis, atom_coords=coords, atom_symbols=list(basis.atom_symbols) + ["X"],
                               shell_centers=centers, shell_params_idx=params)


def _fd(oracle_mod, basis, q, kappa, h=1e-4):
    b = _private_copy(basis, q[0])
    new_atom, new_shell = b.natoms - 1, b.nshells - 1
    qq = (new_shell,) + tuple(q[1:])
    out = []
    for d in range(3):
        vals = []
        for sgn in (1, -1):
            r = b.atom_coords.copy()
            r[new_atom, d] += sgn * h
            vals.append(oracle_mod.Oracle(dataclasses.replace(b, atom_coords=r)).eri(*qq, kappa=kappa))
        out.append((vals[0] - vals[1]) / (2 * h))
    return np.array(out)


def _gh_basis(spher=True):
    coords = np.array([[0.0, 0.2, -0.1], [1.3, 0.9, 0.4], [-1.2, 1.0, 0.5]])
    O = [(0, [0.5], [1.0]), (1, [0.9], [1.0]), (2, [1.2], [1.0]), (3, [0.8], [1.0]), (4, [0.7], [1.0]),
         (2, [3.0, 0.7], [0.5, 0.6])]
    H = [(0, [0.4], [1.0]), (1, [0.6], [1.0]), (3, [1.1], [1.0])]
    return build_basis(["O", "H", "H"], coords, {"O": O, "H": H}, spher_ao=spher)


@pytest.mark.parametrize("spher", [True, False], ids=["spherical", "cartesian"])
@pytest.mark.parametrize("kappa", [0.0, 0.4])
def test_derivative_equals_central_differences(oracle_mod, spher, kappa):
    b = _gh_basis(spher)
    o = oracle_mod.Oracle(b)
    rng = np.random.default_rng(7)
    quartets = [(4, 1, 3, 0), (3, 8, 2, 6), (2, 2, 4, 7), (1, 0, 0, 1), (0, 6, 5, 2), (5, 4, 8, 3), (8, 5, 1, 4)]
    quartets += [tuple(int(x) for x in rng.integers(0, b.nshells, 4)) for _ in range(6)]
    for q in quartets:
        ana = o.eri_deriv(*q, kappa=kappa)
        fd = _fd(oracle_mod, b, q, kappa)
        scale = max(1.0, np.abs(ana).max())
        assert np.abs(ana - fd).max() <= 2e-7 * scale, (q, np.abs(ana - fd).max())


def test_contracted_derivative_equals_central_differences(oracle_mod):
    b = tiny_basis(lmax=3, contracted=True)
    o = oracle_mod.Oracle(b)
    for q in [(3, 4, 0, 6), (4, 3, 5, 1), (6, 3, 3, 2), (2, 4, 7, 5)]:
        ana = o.eri_deriv(*q)
        fd = _fd(oracle_mod, b, q, 0.0)
        assert np.abs(ana - fd).max() <= 2e-7 * max(1.0, np.abs(ana).max()), q


def test_translational_invariance(oracle_mod):
    """sum over the four centres of d(ab|cd)/dX = 0: each centre's
    derivative is the first-shell derivative of a permuted quartet."""
    b = _gh_basis()
    o = oracle_mod.Oracle(b)
    for (a, bb, c, d) in [(4, 1, 3, 0), (3, 8, 2, 6), (5, 4, 8, 3)]:
        da = o.eri_deriv(a, bb, c, d)
        db = o.eri_deriv(bb, a, c, d).transpose(0, 2, 1, 3, 4)
        dc = o.eri_deriv(c, d, a, bb).transpose(0, 3, 4, 1, 2)
        dd = o.eri_deriv(d, c, a, bb).transpose(0, 3, 4, 2, 1)
        tot = da + db + dc + dd
        assert np.abs(tot).max() <= 1e-12 * max(1.0, np.abs(da).max())


@pytest.mark.parametrize("kappa", [0.0, 0.4])
def test_gradient_equals_differences_of_the_energy(oracle_mod, kappa):
    b = tiny_basis(lmax=3, contracted=True)
    D = random_density(b.nao, seed=3, nocc=4)
    g = oracle_mod.Oracle(b).coulomb_grad(D, kappa=kappa)
    fd = fd_gradient(lambda x: oracle_mod.Oracle(x).coulomb_energy(D, kappa=kappa), b)
    assert np.abs(g - fd).max() <= 1e-7 * max(1.0, np.abs(g).max())
    assert np.abs(g.sum(axis=0)).max() <= 1e-10 * max(1.0, np.abs(g).max())


def test_gradient_atom_selection(oracle_mod):
    b = systems.water_system(2, 6.0, "small", seed=5)
    D = random_density(b.nao, seed=2, nocc=5)
    o = oracle_mod.Oracle(b)
    full = o.coulomb_grad(D)
    part = o.coulomb_grad(D, atoms=[1, 4])
    # same sums; only the order of the threads' partial sums differs
    assert np.abs(part[[1, 4]] - full[[1, 4]]).max() <= 1e-13 * np.abs(full).max()
    assert not np.any(np.delete(part, [1, 4], axis=0))

