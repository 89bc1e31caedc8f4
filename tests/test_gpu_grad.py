"""Differentiation form on the GPU: the Coulomb energy E_J of a density and
its nuclear gradient (gi_coulomb_energy_grad), checked against the oracle's
energy, central differences (h = 1e-4 bohr), the translational and
rotational invariances, and through the torch.autograd interface."""
import numpy as np
import pytest

import pyoracle
from gigpu import lib, systems
from gigpu.basis import build_basis
from grad_helpers import fd_gradient, random_density, tiny_basis

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    pyoracle.build()
    c = lib.Context(0)
    yield c
    c.close()


def gpu_energy(ctx, D, kappa=0.0):
    def f(basis):
        ctx.set_basis(basis)
        return ctx.coulomb_energy(D, kappa, grad=False)
    return f


@pytest.mark.parametrize("lmax,contracted,kappa", [(2, True, 0.0), (2, False, 0.0), (3, True, 0.0),
                                                   (2, True, 0.4)])
def test_energy_matches_oracle(ctx, lmax, contracted, kappa):
    basis = tiny_basis(lmax=lmax, contracted=contracted)
    D = random_density(basis.nao, seed=3)
    ctx.set_basis(basis)
    e = ctx.coulomb_energy(D, kappa, grad=False)
    ref = pyoracle.Oracle(basis).coulomb_energy(D, kappa)
    assert abs(e - ref) <= 1e-12 * max(1.0, abs(ref))


def test_energy_water_dimer(ctx):
    basis = systems.water_system(2, 5.0, "small", seed=21)
    D = random_density(basis.nao, seed=4, nocc=10)
    ctx.set_basis(basis)
    e, g = ctx.coulomb_energy(D)
    ref = pyoracle.Oracle(basis).coulomb_energy(D)
    assert abs(e - ref) <= 1e-12 * abs(ref)
    # the energy of the gradient call is the same number
    assert abs(e - ctx.coulomb_energy(D, grad=False)) <= 1e-12 * abs(e)


@pytest.mark.parametrize("contracted,kappa", [(True, 0.0), (False, 0.0), (True, 0.4)])
def test_gradient_vs_oracle_differences(ctx, contracted, kappa):
    basis = tiny_basis(lmax=2, contracted=contracted)
    D = random_density(basis.nao, seed=5, nocc=6)
    ctx.set_basis(basis)
    _, g = ctx.coulomb_energy(D, kappa)
    g_fd = fd_gradient(lambda b: pyoracle.Oracle(b).coulomb_energy(D, kappa), basis)
    assert np.abs(g - g_fd).max() <= 1e-7, np.abs(g - g_fd).max()


@pytest.mark.parametrize("lmax", [2])
def test_cartesian_energy_and_gradient(ctx, lmax):
    """Cartesian AOs (no solid-harmonic transform in the pair tables): energy
    against the oracle, gradient against central differences of the oracle
    (l <= 2: the gradient kernels take pairs with l0 + l1 <= 5)."""
    basis = tiny_basis(lmax=lmax, contracted=True, spher_ao=False)
    D = random_density(basis.nao, seed=7, nocc=6)
    ctx.set_basis(basis)
    e, g = ctx.coulomb_energy(D, 0.0)
    ref = pyoracle.Oracle(basis).coulomb_energy(D, 0.0)
    assert abs(e - ref) <= 1e-12 * max(1.0, abs(ref))
    g_fd = fd_gradient(lambda b: pyoracle.Oracle(b).coulomb_energy(D, 0.0), basis)
    assert np.abs(g - g_fd).max() <= 1e-7, np.abs(g - g_fd).max()


def test_gradient_translation_invariance(ctx):
    basis = systems.water_system(3, 6.0, "small", seed=22)
    D = random_density(basis.nao, seed=6, nocc=15)
    ctx.set_basis(basis)
    _, g = ctx.coulomb_energy(D)
    scale = np.abs(g).max()
    assert scale > 1e-3
    assert np.abs(g.sum(axis=0)).max() <= 1e-10 * scale


def test_gradient_rotation_invariance_s_functions(ctx):
    """With a fixed AO density, p and d functions keep their lab-frame axes
    when the nuclei move, so only an s-only basis makes E_J invariant under
    rotations of the nuclei: zero total torque."""
    from gigpu.basis import build_basis
    rng = np.random.default_rng(23)
    R = rng.uniform(-3.0, 3.0, (5, 3))
    shells = {"H": [(0, [0.3], [1.0]), (0, [1.1], [1.0]), (0, [4.0, 0.9], [0.3, 0.8])]}
    basis = build_basis(["H"] * 5, R, shells)
    D = random_density(basis.nao, seed=10, nocc=4)
    ctx.set_basis(basis)
    _, g = ctx.coulomb_energy(D)
    scale = np.abs(g).max()
    assert np.abs(g.sum(axis=0)).max() <= 1e-10 * scale
    torque = np.cross(basis.atom_coords, g).sum(axis=0)
    assert np.abs(torque).max() <= 1e-10 * scale * np.abs(R).max()


def test_gradient_cluster_vs_differences(ctx):
    """(H2O)8, O 6s4p2d / H 3s1p (320 AOs), D = C_occ C_occ^T with nocc = 40:
    all 72 coordinates against central differences of the GPU energy."""
    basis = systems.config_gradient_cluster()
    D = systems.occupied_density(basis.nao, 40, seed=7)
    ctx.set_basis(basis)
    _, g = ctx.coulomb_energy(D)
    g_fd = fd_gradient(gpu_energy(ctx, D), basis)
    assert np.abs(g - g_fd).max() <= 1e-6, np.abs(g - g_fd).max()


def test_screened_energy_and_gradient(ctx):
    """Density-weighted Schwarz screening (grad_screen = v: shell pairs whose
    bound on every term is below 10^-v are left out) against the unscreened
    result on (H2O)8: measured differences 2.8e-14 (energy) and 6.8e-14
    (gradient) at v = 14, 5.6e-12 and 9.0e-12 at v = 12; the bars allow ten
    times that.  At v = 8 pairs are dropped (the energy moves) and the
    result stays within 1e-6."""
    basis = systems.config_gradient_cluster()
    D = systems.occupied_density(basis.nao, 40, seed=7)
    ctx.set_basis(basis)
    e0, g0 = ctx.coulomb_energy(D)
    out = {}
    try:
        for v in (14, 12, 8):
            ctx.set_option("grad_screen", v)
            out[v] = ctx.coulomb_energy(D)
    finally:
        ctx.set_option("grad_screen", 0)
    for v, (de, dg) in ((14, (3e-13, 7e-13)), (12, (6e-11, 9e-11)), (8, (1e-6, 1e-5))):
        e, g = out[v]
        print("v=%d: dE %.2e max|dg| %.2e" % (v, e - e0, np.abs(g - g0).max()))
        assert abs(e - e0) <= de, v
        assert np.abs(g - g0).max() <= dg, v
    assert out[8][0] != e0


def test_autograd_function(ctx):
    import torch
    from gigpu.autograd import coulomb_energy
    basis = tiny_basis(lmax=2, contracted=True)
    D = random_density(basis.nao, seed=8, nocc=6)
    R = torch.tensor(basis.atom_coords, dtype=torch.float64, requires_grad=True)
    e = coulomb_energy(R, ctx, basis, D)
    e.backward()
    ctx.set_basis(basis)
    e_ref, g_ref = ctx.coulomb_energy(D)
    assert abs(e.item() - e_ref) <= 1e-13 * abs(e_ref)
    assert np.abs(R.grad.numpy() - g_ref).max() <= 1e-13 * np.abs(g_ref).max()
    # a second-order check of the Function through torch: directional derivative
    v = torch.randn(R.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    h = 1e-4
    ep = coulomb_energy((R + h * v).detach(), ctx, basis, D).item()
    em = coulomb_energy((R - h * v).detach(), ctx, basis, D).item()
    assert abs((ep - em) / (2 * h) - float((R.grad * v).sum())) <= 1e-7


def test_gradient_rejects_h_shells(ctx):
    """The gradient's R tables reach order 4 lmax + 1, within the Boys table
    (order 20) for l <= 4; a basis with an h shell gets the energy and a
    clear error for the gradient."""
    coords = np.array([[0.0, 0.1, -0.2], [1.4, 1.1, 0.3]])
    spec = {"O": [(0, [0.45], [1.0]), (5, [0.9], [1.0])], "H": [(0, [0.3], [1.0])]}
    basis = build_basis(["O", "H"], coords, spec)
    D = random_density(basis.nao, seed=9)
    ctx.set_basis(basis)
    e = ctx.coulomb_energy(D, grad=False)
    assert np.isfinite(e) and e > 0
    with pytest.raises(lib.GigpuError, match="l <= 4"):
        ctx.coulomb_energy(D)
