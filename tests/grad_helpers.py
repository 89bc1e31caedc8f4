"""Shared inputs of the Coulomb energy / nuclear gradient tests."""
import dataclasses

import numpy as np

from gigpu.basis import build_basis


def tiny_basis(lmax=2, contracted=True, spher_ao=True):
    """O + H + H with mixed shells (s, p, d; f when lmax = 3) and, when
    contracted, a 3-primitive s and a 2-primitive p shell on O."""
    coords = np.array([[0.0, 0.1, -0.2], [1.4, 1.1, 0.3], [-1.5, 0.9, 0.4]])
    O = [(0, [0.45], [1.0]), (1, [0.8], [1.0]), (2, [1.1], [1.0])]
    if contracted:
        O += [(0, [6.0, 1.3, 0.35], [0.15, 0.55, 0.45]), (1, [3.0, 0.6], [0.4, 0.7])]
    else:
        O += [(0, [2.0], [1.0])]
    if lmax >= 3:
        O += [(3, [0.9], [1.0])]
    H = [(0, [0.3], [1.0]), (0, [1.5], [1.0]), (1, [0.75], [1.0])]
    return build_basis(["O", "H", "H"], coords, {"O": O, "H": H}, spher_ao=spher_ao)


def moved(basis, atom, d, h):
    r = basis.atom_coords.copy()
    r[atom, d] += h
    return dataclasses.replace(basis, atom_coords=r)


def fd_gradient(energy, basis, h=1e-4):
    """Central differences of energy(basis) over every atom coordinate."""
    g = np.zeros((basis.natoms, 3))
    for a in range(basis.natoms):
        for d in range(3):
            g[a, d] = (energy(moved(basis, a, d, h)) - energy(moved(basis, a, d, -h))) / (2 * h)
    return g


def random_density(nao, seed, nocc=None):
    """Symmetric density: C_occ C_occ^T of a random orthonormal C, or a
    random symmetric matrix when nocc is None."""
    rng = np.random.default_rng(seed)
    if nocc is None:
        a = rng.standard_normal((nao, nao))
        return 0.5 * (a + a.T)
    q, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    return q[:, :nocc] @ q[:, :nocc].T
