"""Properties of the decomposition at the benchmark size ((H2O)16, O
8s5p3d1f / H 4s2p1d, 1200 AOs, tau = 1e-6), where the oracle would take
hours: the reconstruction bound, positive semidefinite residual, and the MO
transform's symmetry and linearity."""
import numpy as np
import pytest

from gigpu import lib, systems

pytestmark = pytest.mark.gpu
TAU = 1e-6


@pytest.fixture(scope="module")
def large():
    basis = systems.config_large_cluster()
    ctx = lib.Context(0)
    ctx.set_basis(basis)
    yield ctx, basis
    ctx.close()


def test_reconstruction_bound_at_full_size(large):
    ctx, basis = large
    ctx.chol_build(tau=TAU, flags=lib.CHOL_DECODE_FIXED)
    n_fixed = ctx.chol_dims()["nvecs"]
    mx, mn = ctx.chol_residual()
    # every residual diagonal of the Step-1 pivot set is below tau; V - R^T R is PSD
    assert mx <= TAU * (1 + 1e-9)
    assert mn >= -1e-10
    ctx.chol_build(tau=TAU)
    assert ctx.chol_dims()["nvecs"] == n_fixed
    assert 4 * basis.nao <= n_fixed <= 15 * basis.nao   # the cap of TwoStepCholesky.f90:119,155


def test_mo_transform_symmetry_and_linearity_at_full_size(large):
    ctx, basis = large
    ctx.chol_build(tau=TAU)
    rng = np.random.default_rng(11)
    Cm, _ = np.linalg.qr(rng.normal(size=(basis.nao, basis.nao)))
    n = 40
    R = ctx.mo_transform(Cm, 1, n, Cm, 1, n)                # (nvecs, n*n), a fastest
    nv = R.shape[0]
    R3 = R.reshape(nv, n, n)                                  # [k, b, a]
    scale = np.abs(R3).max()
    assert np.abs(R3 - R3.transpose(0, 2, 1)).max() <= 1e-12 * scale
    # linear in CB
    C1 = Cm[:, :n]
    C2 = rng.normal(size=(basis.nao, n))
    CA = Cm.copy()
    Ra = ctx.mo_transform(CA, 1, 8, np.hstack([C1, C2, C1 + C2]), 1, 3 * n).reshape(nv, 3 * n, 8)
    assert np.abs(Ra[:, 2 * n:] - Ra[:, :n] - Ra[:, n:2 * n]).max() <= 1e-11 * np.abs(Ra).max()
    # the split-BF16 form stays within FP32 accumulation error of the FP64 transform
    R16 = ctx.mo_transform(Cm, 1, n, Cm, 1, n, flags=lib.MO_SPLIT_BF16)
    assert np.abs(R16 - R).max() <= 1e-6 * scale
