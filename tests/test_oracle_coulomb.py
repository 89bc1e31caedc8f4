"""The oracle's Coulomb energy E_J = 1/2 sum D_pq (pq|rs) D_rs (the checker
of the differentiation form), pinned by an independent contraction and a
closed form."""
import math

import numpy as np
import pytest

import pyoracle
from gigpu.basis import build_basis
from grad_helpers import random_density, tiny_basis


@pytest.fixture(scope="module", autouse=True)
def _built():
    pyoracle.build()


def dense_eri(o, basis):
    nao = basis.nao
    loc = basis.shell_loc() - 1
    n = [basis.shell_nang(s) for s in range(basis.nshells)]
    G = np.zeros((nao,) * 4)
    for a in range(basis.nshells):
        for b in range(basis.nshells):
            for c in range(basis.nshells):
                for d in range(basis.nshells):
                    G[loc[a]:loc[a] + n[a], loc[b]:loc[b] + n[b], loc[c]:loc[c] + n[c],
                      loc[d]:loc[d] + n[d]] = o.eri(a, b, c, d)
    return G


@pytest.mark.parametrize("lmax,kappa", [(2, 0.0), (3, 0.0), (2, 0.4)])
def test_energy_equals_unsymmetrized_contraction(lmax, kappa):
    basis = tiny_basis(lmax=lmax, contracted=True)
    o = pyoracle.Oracle(basis)
    D = random_density(basis.nao, seed=11)
    if kappa == 0.0:
        G = dense_eri(o, basis)
        ref = 0.5 * np.einsum("pq,pqrs,rs->", D, G, D)
        assert abs(o.coulomb_energy(D) - ref) <= 1e-12 * max(1.0, abs(ref))
    else:
        # attenuated operator: only the symmetry-reduced sum is at hand, so
        # compare with the unattenuated energy, which must be larger for a
        # positive semidefinite density
        Dp = random_density(basis.nao, seed=12, nocc=5)
        assert 0.0 < o.coulomb_energy(Dp, kappa) < o.coulomb_energy(Dp)


def test_single_s_function_closed_form():
    alpha = 0.7
    basis = build_basis(["H"], np.zeros((1, 3)), {"H": [(0, [alpha], [1.0])]})
    o = pyoracle.Oracle(basis)
    # (ss|ss) = 2 sqrt(alpha/pi) for a normalized s function; the Boys table
    # is accurate to 4.3e-12 relative (boys.f90:30)
    ref = math.sqrt(alpha / math.pi)
    assert abs(o.coulomb_energy(np.ones((1, 1))) - ref) <= 4.3e-12 * ref
