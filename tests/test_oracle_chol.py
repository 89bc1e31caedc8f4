"""Pins of the oracle's two-step Cholesky decomposition by the mathematics it
must satisfy, independent of any reference output:

* the vectors reproduce the factorized columns of V = (pq|rs) exactly,
  R^T R (:, P) = V(:, P) (Cholesky on the Step-2 pivot set P);
* every element of V - R^T R is within tau (|E_pq| <= sqrt(D_p D_q) with the
  residual diagonal D <= tau);
* Inv_L is the inverse Cholesky factor of V(P, P);
* the pivot set is a function of V alone: splitting the orbital pairs into
  subsets (MaxBlockDim) changes storage, not the decomposition.
Dense V comes from the oracle's ERIs, which are pinned separately against an
independent Obara-Saika evaluation."""
import numpy as np
import pytest

from chol_helpers import cholesky_factor, dense_v, step2_columns
from gigpu import systems


def ref_meta(r):
    return dict(shell_pairs=r.shell_pairs, shell_pair_loc=r.shell_pair_loc, pivot_shell_pairs=r.pivot_shell_pairs,
                pivot_shell_pair_loc=r.pivot_shell_pair_loc, pivot_shell_pair_dim=r.pivot_shell_pair_dim,
                pivot_orb_pairs=r.pivot_orb_pairs)


@pytest.fixture(scope="module", params=["uncontracted", "contracted", "contracted-cartesian"])
def system(oracle_mod, request):
    contracted = request.param != "uncontracted"
    cart = request.param == "contracted-cartesian"
    b = systems.water_system(1, 6.0, "small", seed=7, contracted=contracted, spher_ao=not cart)
    o = oracle_mod.Oracle(b)
    # Cartesian d shells: the reference's Step-2 decode fails (see
    # test_cartesian_same_shell_pairs_are_dependent); the Step-1 decode is used.
    flags = 1 if cart else 0
    r = o.chol(tau=1e-6, flags=flags)
    V = dense_v(o, b, r.shell_pairs)
    return b, o, r, V, flags


@pytest.mark.parametrize("flags", [0, 1])
def test_vectors_reproduce_pivot_columns_and_stay_within_tau(oracle_mod, system, flags):
    b, o, r0, V, sys_flags = system
    r = o.chol(tau=1e-6, flags=flags)
    if flags != sys_flags and not b.spher_ao:
        assert r.status == 40 and "info=" in r.message
        return
    assert r.status == 0 and r.nvecs > 0
    R = r.full_vectors()
    E = V - R.T @ R
    P = step2_columns(b, ref_meta(r))
    assert len(np.unique(P)) == r.nvecs
    vmax = np.abs(V).max()
    assert np.abs(E[:, P]).max() <= 1e-12 * vmax
    if flags == 1:
        # the tau bound holds for the pivots Step 1 selected; the reference's
        # decode factorizes a different same-shell set and can exceed it
        assert np.abs(E).max() <= 1e-6 * (1 + 1e-9)
    assert np.diag(E).min() >= -1e-12 * vmax          # residual diagonal is non-negative
    # Inv_L is the inverse Cholesky factor of V(P, P)
    L = cholesky_factor(r.inv_l)
    assert np.abs(L @ L.T - V[np.ix_(P, P)]).max() <= 1e-11 * vmax
    assert np.allclose(np.triu(r.inv_l, 1), 0.0)
    assert np.all(np.diag(r.inv_l) > 0)


def test_fixed_decode_factorizes_step1_pivots(oracle_mod, system):
    """With the Step-1 storage-order decode, Step 2 factorizes exactly the
    pivots Step 1 selected."""
    b, o, _, V, _ = system
    r = o.chol(tau=1e-6, flags=1)
    P = step2_columns(b, ref_meta(r))
    assert np.array_equal(np.sort(P), np.sort(r.pivot_order - 1))


def test_subsets_do_not_change_the_decomposition(oracle_mod, system):
    b, o, r0, V, flags = system
    r1 = o.chol(tau=1e-6, max_block_dim=200, flags=flags)
    assert r1.nsubsets > 1
    assert np.array_equal(r1.pivot_order, r0.pivot_order)
    R0, R1 = r0.full_vectors(), r1.full_vectors()
    assert np.abs(R0 - R1).max() <= 1e-12 * max(1.0, np.abs(R0).max())


def test_residual_bounds_pivot_count(oracle_mod, system):
    """A tighter threshold keeps more vectors and a smaller residual."""
    b, o, _, V, _ = system
    ra, rb = o.chol(tau=1e-4, flags=1), o.chol(tau=1e-8, flags=1)
    assert ra.nvecs < rb.nvecs
    for r, tau in ((ra, 1e-4), (rb, 1e-8)):
        R = r.full_vectors()
        assert np.abs(np.diag(V - R.T @ R)).max() <= tau * (1 + 1e-9)


def test_cartesian_same_shell_pairs_are_dependent(oracle_mod):
    """Why the reference's Step-2 decode fails on Cartesian d shells: on one
    shell, d_xx d_yy and d_xy d_xy are the same function x^2 y^2 up to the
    component norms (ratio 1/3), so their Coulomb columns are proportional.
    Step 1 never selects both (the second one's residual vanishes), but the
    chol2_pq2p_ge_q decode (TwoStepCholesky_Step2.f90:422-466) can map a
    selected same-shell pivot onto its dependent partner, and the pivot block
    V(P, P) is then singular: the oracle reports status 40 with the message
    of the reference's failed factorization (real_linalg.f90:54)."""
    b = systems.water_system(1, 6.0, "small", seed=7, spher_ao=False)
    o = oracle_mod.Oracle(b)
    d = [i for i in range(b.nshells) if b.shell_l(i) == 2][0]
    xx, xy, yy = 0, 1, 3
    for s in range(0, b.nshells, 3):
        for t in range(0, s + 1, 2):
            G = o.eri(d, d, s, t)
            assert np.allclose(G[xx, yy], G[xy, xy] / 3.0, rtol=1e-13, atol=1e-15)
    assert o.chol(tau=1e-6).status == 40
    assert o.chol(tau=1e-6, flags=1).status == 0


@pytest.mark.parametrize("n,basis", [(2, "small"), (1, "large")])
def test_debug_accuracy_converges_below_tau(oracle_mod, n, basis):
    """tau = 1e-10 (CHOL_ACCURACY_DEBUG) is below the residual quantum the
    orderings would use at 2^-32 of MaxDrs; the thresholds compare raw values,
    so the build converges with max residual <= tau, and the quantized and raw
    orderings keep the same number of vectors."""
    b = systems.water_system(n, 6.0, basis, seed=1)
    o = oracle_mod.Oracle(b)
    q = o.chol(tau=1e-10, flags=1 | 2)
    r = o.chol(tau=1e-10, flags=1 | 2 | 4)
    for x in (q, r):
        assert x.status == 0, x.message
        assert 0.0 < x.max_residual <= 1e-10
    assert q.nvecs == r.nvecs
    assert q.nvecs < 15 * b.nao
