"""GPU parity of the two-step Cholesky decomposition against the oracle.

Exact: pivot sequence and count, every TChol2Vecs index array.  To ERI
accuracy: Vdiag (1e-12), the columns of V the vectors reproduce,
R^T R (:, P) = V(:, P), and the factorized pivot block L L^T = V(P, P).
Inv_L and the vectors themselves are ill-conditioned functions of V (the
GPU evaluates the Boys function once per primitive quartet at top order L
where the reference does it per (e0|f0) block, which moves ERIs by up to
~1e-13); their elementwise difference is bounded by a relative ERI
perturbation of 1e-14 amplified by cond(L)^2."""
import numpy as np
import pytest

from chol_helpers import cholesky_factor, dense_v, step2_columns
from gigpu import gammcor, systems
from pyoracle import full_vectors

pytestmark = pytest.mark.gpu


def _ref_meta(r):
    return dict(shell_pairs=r.shell_pairs, shell_pair_loc=r.shell_pair_loc, pivot_shell_pairs=r.pivot_shell_pairs,
                pivot_shell_pair_loc=r.pivot_shell_pair_loc, pivot_shell_pair_dim=r.pivot_shell_pair_dim,
                pivot_orb_pairs=r.pivot_orb_pairs)


def _compare(ctx, o, b, tau, flags=0, max_block_dim=4000, oracle_flags=None):
    ref = o.chol(tau=tau, flags=flags if oracle_flags is None else oracle_flags, max_block_dim=max_block_dim)
    assert ref.status == 0, ref.message
    ctx.set_basis(b)
    ctx.chol_build(tau=tau, flags=flags, max_block_dim=max_block_dim)
    m = ctx.chol_meta()
    assert m["nvecs"] == ref.nvecs
    assert m["norb_pairs"] == ref.norb_pairs and m["nshell_pairs"] == ref.nshell_pairs
    assert m["nsubsets"] == ref.nsubsets
    for k in ("shell_pairs", "shell_pair_loc", "shell_pair_dim", "subset_dim", "subset_bounds", "pivot_order",
              "pivots", "pivot_orb_pairs", "pivot_shell_pairs", "pivot_shell_pair_loc", "pivot_shell_pair_dim"):
        assert np.array_equal(m[k], getattr(ref, k)), k
    assert np.abs(ctx.chol_vdiag() - ref.vdiag).max() <= 1e-12
    # vectors: reproduced columns agree to ERI accuracy
    R = ctx.chol_rkpq()
    Rg = full_vectors(m["shell_pair_loc"], m["shell_pair_dim"], R, m["norb_pairs"])
    Rr = ref.full_vectors()
    P = step2_columns(b, m)
    Gg, Gr = Rg.T @ Rg[:, P], Rr.T @ Rr[:, P]
    gmax = np.abs(Gr).max()
    assert np.abs(Gg - Gr).max() <= 1e-11 * gmax
    # Inv_L: the factorized block agrees to ERI accuracy
    Lg_inv, Lr_inv = ctx.chol_inv_l(), ref.inv_l
    Lg, Lr = cholesky_factor(Lg_inv), cholesky_factor(Lr_inv)
    assert np.abs(Lg @ Lg.T - Lr @ Lr.T).max() <= 1e-11 * gmax
    # elementwise, bounded by the conditioning of L
    cond = np.linalg.norm(Lr, 2) * np.linalg.norm(Lr_inv, 2)
    assert np.abs(Lg_inv - Lr_inv).max() <= 1e-14 * cond ** 2 * np.abs(Lr_inv).max()
    assert np.abs(Rg - Rr).max() <= 1e-14 * cond ** 2 * max(1.0, np.abs(Rr).max())
    return ref, m, R


@pytest.mark.parametrize("n,basis,seed", [(1, "small", 1), (2, "small", 2), (1, "large", 3)])
def test_chol_matches_oracle(gpu_ctx, oracle_mod, n, basis, seed):
    b = systems.water_system(n, 6.0, basis, seed=seed)
    _compare(gpu_ctx, oracle_mod.Oracle(b), b, 1e-6)


def test_chol_small_subsets_and_fixed_decode(gpu_ctx, oracle_mod):
    b = systems.water_system(2, 6.0, "small", seed=5)
    ref, m, _ = _compare(gpu_ctx, oracle_mod.Oracle(b), b, 1e-6, flags=1, max_block_dim=700)
    assert m["nsubsets"] == ref.nsubsets > 1


def test_reconstruction_error_within_tau(gpu_ctx, oracle_mod):
    """V - R^T R: diagonal residuals within tau for the fixed decode, and the
    full matrix within tau (Cholesky error bound: |V_pq - (R^T R)_pq| <= sqrt(D_p D_q))."""
    b = systems.water_system(1, 6.0, "small", seed=7)
    o = oracle_mod.Oracle(b)
    gpu_ctx.set_basis(b)
    gpu_ctx.chol_build(tau=1e-6, flags=1)
    m = gpu_ctx.chol_meta()
    R = gpu_ctx.chol_rkpq()
    Rf = full_vectors(m["shell_pair_loc"], m["shell_pair_dim"], R, m["norb_pairs"])
    V = dense_v(o, b, m["shell_pairs"])
    E = V - Rf.T @ Rf
    assert np.abs(np.diag(E)).max() <= 1e-6 * (1 + 1e-6)
    assert np.abs(E).max() <= 1e-6 * (1 + 1e-6)


def test_gammcor_interface_and_mo(gpu_ctx, oracle_mod):
    b = systems.water_system(2, 6.0, "small", seed=9)
    vecs = gammcor.chol_gammcor_Rkpq(b, gammcor.CHOL_ACCURACY_TIGHT, ctx=gpu_ctx)
    o = oracle_mod.Oracle(b)
    ref = o.chol(tau=1e-6)
    assert vecs.Chol2Data.NVecs == ref.nvecs
    rng = np.random.default_rng(0)
    C, _ = np.linalg.qr(rng.normal(size=(b.nao, b.nao)))
    C = np.asfortranarray(C)
    nocc, nvirt = 10, 30
    Rab = gammcor.chol_gammcor_Rkab(None, C, 1, nocc, C, nocc + 1, nocc + nvirt, 100000, vecs, b,
                                    gammcor.ORBITAL_ORDERING_DALTON)
    # the MO transform is checked on the GPU's own vectors
    ref.set_rkpq(vecs.rkpq())
    Rref = ref.mo_transform(C, 1, nocc, C, nocc + 1, nocc + nvirt)
    assert np.abs(Rab - Rref).max() <= 1e-10 * max(1.0, np.abs(Rref).max())
    # a0 > 1 keeps the reference's column mask
    Rab2 = gammcor.chol_gammcor_Rkab(None, C, 3, 12, C, 1, 5, 100000, vecs, b, gammcor.ORBITAL_ORDERING_DALTON)
    Rref2 = ref.mo_transform(C, 3, 12, C, 1, 5)
    assert np.abs(Rab2 - Rref2).max() <= 1e-10 * max(1.0, np.abs(Rref2).max())
    with pytest.raises(gammcor.GammcorError, match="Buffer size too small"):
        gammcor.chol_gammcor_Rkab(None, C, 1, nocc, C, 1, 5, 0, vecs, b, gammcor.ORBITAL_ORDERING_DALTON)


def test_range_separated_cholesky(gpu_ctx, oracle_mod):
    """Omega > 0 (Kappa = 1/Omega^2, Cholesky_Gammcor.f90:411-415): the
    long-range Coulomb matrix decomposed with the same pivots as the oracle."""
    b = systems.water_system(1, 6.0, "small", seed=13)
    o = oracle_mod.Oracle(b)
    omega = 0.7
    kappa = 1.0 / omega ** 2
    ref = o.chol(tau=1e-6, kappa=kappa)
    assert ref.status == 0, ref.message
    gpu_ctx.set_basis(b)
    gpu_ctx.chol_build(tau=1e-6, kappa=kappa)
    m = gpu_ctx.chol_meta()
    assert m["nvecs"] == ref.nvecs
    assert np.array_equal(m["pivot_order"], ref.pivot_order)
    assert np.abs(gpu_ctx.chol_vdiag() - ref.vdiag).max() <= 1e-12
    Rg = full_vectors(m["shell_pair_loc"], m["shell_pair_dim"], gpu_ctx.chol_rkpq(), m["norb_pairs"])
    Rr = ref.full_vectors()
    P = step2_columns(b, m)
    Gg, Gr = Rg.T @ Rg[:, P], Rr.T @ Rr[:, P]
    assert np.abs(Gg - Gr).max() <= 1e-11 * np.abs(Gr).max()
    # the attenuated operator has fewer significant vectors than the full one
    full = o.chol(tau=1e-6)
    assert ref.nvecs < full.nvecs
    vecs = gammcor.chol_gammcor_Rkpq(b, gammcor.CHOL_ACCURACY_TIGHT, Omega=omega, ctx=gpu_ctx)
    assert vecs.Chol2Data.NVecs == ref.nvecs


@pytest.mark.parametrize("n,basis", [(2, "large"), (4, "large")])
def test_micro_iterations_one_launch(gpu_ctx, n, basis):
    """The micro iterations of every macro iteration in one launch (panels and
    their MFMA rank updates of the batch block from LDS) against a panel and
    an update launch per 16 steps: the same pivots, hence the same Step-2
    factorization and vectors."""
    b = systems.water_system(n, 8.0, basis, seed=31)
    gpu_ctx.set_basis(b)
    res = {}
    try:
        for m in (0, 1024):
            gpu_ctx.set_option("micro_persist_max", m)
            gpu_ctx.chol_build(tau=1e-6)
            res[m] = (np.array(gpu_ctx.chol_meta()["pivot_order"]), np.array(gpu_ctx.chol_rkpq()),
                      gpu_ctx.chol_residual()[0])
    finally:
        gpu_ctx.set_option("micro_persist_max", 512)
    assert np.array_equal(res[0][0], res[1024][0])
    diff = float(np.abs(res[0][1] - res[1024][1]).max())
    print("n=%d: %d vectors, max |R(one launch) - R(panels)| = %.2e, residual %.3e / %.3e"
          % (n, len(res[0][0]), diff, res[0][2], res[1024][2]))
    assert diff <= 1e-13 * float(np.abs(res[0][1]).max())
    assert res[1024][2] == res[0][2]
