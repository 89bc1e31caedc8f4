"""The GPU nuclear gradient of E_J against the FP64 derivative-ERI oracle
(or_coulomb_grad: d/dA of the Cartesian Gaussians as raised and lowered
shells, an independent route to the gradient; pinned by central
differences in test_oracle_deriv.py), to 1e-8 Eh/bohr:

* contracted s-f shells, Kappa = 0 and 0.4;
* g shells on O (the l <= 4 limit of the gradient kernels);
* (H2O)1 with O 8s5p3d1f / H 4s2p1d (f-f pairs, l0 + l1 = 6);
* (H2O)8 with O 6s4p2d / H 3s1p (320 AOs, D = C_occ C_occ^T, nocc = 40):
  the nine coordinates of the first water (the oracle's cost grows with the
  number of atoms selected)."""
import numpy as np
import pytest

import pyoracle
from gigpu import lib, systems
from gigpu.basis import build_basis
from grad_helpers import random_density, tiny_basis

pytestmark = pytest.mark.gpu
TOL = 1e-8


@pytest.fixture(scope="module")
def ctx():
    pyoracle.build()
    c = lib.Context(0)
    yield c
    c.close()


def _gpu_grad(ctx, basis, D, kappa=0.0):
    ctx.set_basis(basis)
    return ctx.coulomb_energy(D, kappa)[1]


@pytest.mark.parametrize("kappa", [0.0, 0.4])
def test_contracted_f_shells(ctx, kappa):
    b = tiny_basis(lmax=3, contracted=True)
    D = random_density(b.nao, seed=11, nocc=5)
    g = _gpu_grad(ctx, b, D, kappa)
    ref = pyoracle.Oracle(b).coulomb_grad(D, kappa=kappa)
    assert np.abs(g - ref).max() <= TOL, np.abs(g - ref).max()


@pytest.mark.parametrize("spher", [True, False], ids=["spherical", "cartesian"])
def test_g_shells(ctx, spher):
    coords = np.array([[0.0, 0.2, -0.1], [1.3, 0.9, 0.4], [-1.2, 1.0, 0.5]])
    O = [(0, [0.5], [1.0]), (1, [0.9], [1.0]), (2, [1.2], [1.0]), (3, [0.8], [1.0]), (4, [0.7], [1.0]),
         (4, [2.0, 0.6], [0.4, 0.7])]
    H = [(0, [0.4], [1.0]), (1, [0.6], [1.0]), (2, [1.1], [1.0])]
    b = build_basis(["O", "H", "H"], coords, {"O": O, "H": H}, spher_ao=spher)
    D = random_density(b.nao, seed=12, nocc=6)
    g = _gpu_grad(ctx, b, D)
    ref = pyoracle.Oracle(b).coulomb_grad(D)
    assert np.abs(g - ref).max() <= TOL, np.abs(g - ref).max()


def test_water_large_basis_with_f_pairs(ctx):
    b = systems.water_system(1, 6.0, "large", seed=13)
    D = random_density(b.nao, seed=13, nocc=5)
    g = _gpu_grad(ctx, b, D)
    ref = pyoracle.Oracle(b).coulomb_grad(D)
    assert np.abs(g - ref).max() <= TOL, np.abs(g - ref).max()


def test_water_octamer_first_water(ctx):
    b = systems.config_gradient_cluster()
    assert b.nao == 320 and b.natoms == 24
    D = systems.occupied_density(b.nao, 40, seed=7)
    g = _gpu_grad(ctx, b, D)
    atoms = [0, 1, 2]
    ref = pyoracle.Oracle(b).coulomb_grad(D, atoms=atoms)
    err = np.abs(g[atoms] - ref[atoms]).max()
    print("(H2O)8, first water: max |gpu - oracle| = %.2e Eh/bohr" % err)
    assert err <= TOL
