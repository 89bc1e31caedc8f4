"""GPU parity of the two-step Cholesky decomposition against the oracle:
identical pivot sequence and count, Vdiag to 1e-12, Inv_L / Rkpq / MO
vectors to round-off, and the decomposition's own invariants."""
import numpy as np
import pytest

from gigpu import gammcor, systems

pytestmark = pytest.mark.gpu


def _compare(ctx, o, b, tau, flags=0):
    ref = o.chol(tau=tau, flags=flags)
    assert ref.status == 0, ref.message
    ctx.set_basis(b)
    ctx.chol_build(tau=tau, flags=flags)
    m = ctx.chol_meta()
    assert m["nvecs"] == ref.nvecs
    assert m["norb_pairs"] == ref.norb_pairs and m["nshell_pairs"] == ref.nshell_pairs
    assert np.array_equal(m["shell_pairs"], ref.shell_pairs)
    assert np.array_equal(m["shell_pair_loc"], ref.shell_pair_loc)
    assert np.array_equal(m["subset_bounds"], ref.subset_bounds)
    assert np.array_equal(m["pivot_order"], ref.pivot_order)
    assert np.array_equal(m["pivots"], ref.pivots)
    assert np.array_equal(m["pivot_orb_pairs"], ref.pivot_orb_pairs)
    assert np.abs(ctx.chol_vdiag() - ref.vdiag).max() <= 1e-12
    L = ctx.chol_inv_l()
    assert np.abs(L - ref.inv_l).max() <= 1e-8 * max(1.0, np.abs(ref.inv_l).max())
    R = ctx.chol_rkpq()
    assert np.abs(R - ref.rkpq).max() <= 1e-9 * max(1.0, np.abs(ref.rkpq).max())
    return ref, m, R


@pytest.mark.parametrize("n,basis,seed", [(1, "small", 1), (2, "small", 2), (1, "large", 3)])
def test_chol_matches_oracle(gpu_ctx, oracle_mod, n, basis, seed):
    b = systems.water_system(n, 6.0, basis, seed=seed)
    _compare(gpu_ctx, oracle_mod.Oracle(b), b, 1e-6)


def test_chol_small_subsets_and_fixed_decode(gpu_ctx, oracle_mod):
    b = systems.water_system(2, 6.0, "small", seed=5)
    o = oracle_mod.Oracle(b)
    ref = o.chol(tau=1e-6, max_block_dim=700, flags=1)
    gpu_ctx.set_basis(b)
    gpu_ctx.chol_build(tau=1e-6, max_block_dim=700, flags=1)
    m = gpu_ctx.chol_meta()
    assert m["nsubsets"] == ref.nsubsets > 1
    assert np.array_equal(m["pivot_order"], ref.pivot_order)
    assert np.abs(gpu_ctx.chol_rkpq() - ref.rkpq).max() <= 1e-9 * max(1.0, np.abs(ref.rkpq).max())


def _dense_v(o, b, m):
    """Dense V over BASE orbital pairs in FULL storage order."""
    sp = m["shell_pairs"]
    rows = []
    for s, (A, B) in enumerate(sp):
        na, nb = b.shell_nang(A - 1), b.shell_nang(B - 1)
        for bb in range(nb):
            for a in range(bb if A == B else 0, na):
                rows.append((s, a, bb))
    n = len(rows)
    V = np.zeros((n, n))
    cache = {}
    for i, (s, a, bb) in enumerate(rows):
        for j, (t, c, d) in enumerate(rows):
            key = (s, t)
            if key not in cache:
                A, B = sp[s]
                C_, D = sp[t]
                cache[key] = o.eri(A - 1, B - 1, C_ - 1, D - 1)
            V[i, j] = cache[key][a, bb, c, d]
    return V


def test_reconstruction_error_within_tau(gpu_ctx, oracle_mod):
    """V - R^T R: diagonal residuals within tau for the fixed decode, and the
    full matrix within tau (Cholesky error bound: |V_pq - (R^T R)_pq| <= sqrt(D_p D_q))."""
    b = systems.water_system(1, 6.0, "small", seed=7)
    o = oracle_mod.Oracle(b)
    gpu_ctx.set_basis(b)
    gpu_ctx.chol_build(tau=1e-6, flags=1)
    m = gpu_ctx.chol_meta()
    R = gpu_ctx.chol_rkpq()
    # columns of Rkpq in FULL order (single subset here)
    assert m["nsubsets"] == 1
    Rf = R[0, :m["norb_pairs"], :].T
    V = _dense_v(o, b, m)
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
    Rref = ref.mo_transform(C, 1, nocc, C, nocc + 1, nocc + nvirt)
    assert np.abs(Rab - Rref).max() <= 1e-10 * max(1.0, np.abs(Rref).max())
    # a0 > 1 keeps the reference's column mask
    Rab2 = gammcor.chol_gammcor_Rkab(None, C, 3, 12, C, 1, 5, 100000, vecs, b, gammcor.ORBITAL_ORDERING_DALTON)
    Rref2 = ref.mo_transform(C, 3, 12, C, 1, 5)
    assert np.abs(Rab2 - Rref2).max() <= 1e-10 * max(1.0, np.abs(Rref2).max())
    with pytest.raises(gammcor.GammcorError, match="Buffer size too small"):
        gammcor.chol_gammcor_Rkab(None, C, 1, nocc, C, 1, 5, 0, vecs, b, gammcor.ORBITAL_ORDERING_DALTON)
