"""One-electron integrals on the GPU (gi_one_electron, gi_h0) against the
oracle's restatement of OneElectronInts.f90."""
import numpy as np
import pytest

import pyoracle
from gigpu import lib, systems
from grad_helpers import tiny_basis

pytestmark = pytest.mark.gpu
Z = {"O": 8.0, "H": 1.0, "A": 0.0, "B": 0.0}


@pytest.fixture(scope="module")
def ctx():
    pyoracle.build()
    c = lib.Context(0)
    yield c
    c.close()


def charges(basis):
    return np.array([Z[s] for s in basis.atom_symbols])


def compare(ctx, basis, q, pc=None):
    ctx.set_basis(basis)
    S, T, V = ctx.one_electron(q, pc)
    So, To, Vo = pyoracle.Oracle(basis).one_electron(q, pc)
    for X, Xo in ((S, So), (T, To), (V, Vo)):
        assert np.abs(X - Xo).max() <= 1e-12 * max(1.0, np.abs(Xo).max())
        assert np.array_equal(X, X.T)
    H = ctx.h0(q, pc)
    assert np.abs(H - (To + Vo)).max() <= 1e-12 * max(1.0, np.abs(To + Vo).max())
    return S, T, V


def test_water_cluster_f_shells_with_point_charges(ctx):
    basis = systems.water_system(3, 6.0, "large", seed=31)
    pc = np.array([[3.0, 1.0, -2.0, -0.8], [-1.5, 4.0, 0.5, 0.4]])
    S, _, _ = compare(ctx, basis, charges(basis), pc)
    assert np.abs(np.diag(S) - 1.0).max() <= 1e-12


def test_contracted_shells(ctx):
    basis = tiny_basis(lmax=3, contracted=True)
    compare(ctx, basis, np.array([8.0, 1.0, 1.0]))


def test_cartesian_aos(ctx):
    basis = systems.water_system(2, 5.0, "small", seed=32, spher_ao=False)
    compare(ctx, basis, charges(basis))


def test_dummy_atoms_carry_no_charge(ctx):
    basis = systems.water_system(1, 4.0, "small", seed=33)
    ctx.set_basis(basis)
    _, _, V0 = ctx.one_electron(np.zeros(basis.natoms))
    assert np.abs(V0).max() == 0.0
    _, T1, V1 = ctx.one_electron(None)
    assert np.abs(V1).max() == 0.0 and np.abs(T1).max() > 0.0


def test_gammcor_module_mirror(ctx):
    from gigpu import gammcor, oneelectron
    basis = systems.water_system(2, 5.0, "small", seed=34)
    system = oneelectron.system_of(basis)
    assert list(system.atom_charges()) == [8.0, 1.0, 1.0] * 2
    H = oneelectron.ints1e_gammcor_H0(None, basis, system, ctx)
    _, To, Vo = pyoracle.Oracle(basis).one_electron(system.atom_charges())
    assert np.abs(H - (To + Vo)).max() <= 1e-12 * np.abs(To + Vo).max()
    rng = np.random.default_rng(5)
    C, _ = np.linalg.qr(rng.normal(size=(basis.nao, basis.nao)))
    Hmo = oneelectron.ints1e_gammcor_H0_mo(None, C[:, :12], basis, system, gammcor.ORBITAL_ORDERING_DALTON, ctx)
    assert np.abs(Hmo - C[:, :12].T @ H @ C[:, :12]).max() <= 1e-12 * np.abs(Hmo).max()
    with pytest.raises(gammcor.GammcorError, match="Invalid dimension of matrix C"):
        oneelectron.ints1e_gammcor_H0_mo(None, C[:-1], basis, system, gammcor.ORBITAL_ORDERING_DALTON, ctx)
    # the reference's AO map rejects its own Auto2e ordering
    # (auto2e_interface_AOMap, Auto2eInterface.f90:439)
    with pytest.raises(gammcor.GammcorError, match="Invalid external orbital ordering"):
        oneelectron.ints1e_gammcor_H0_extao(None, basis, system, gammcor.ORBITAL_ORDERING_AUTO2E, ctx)
