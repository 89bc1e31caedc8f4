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


@pytest.fixture(scope="module")
def system(oracle_mod):
    b = systems.water_system(1, 6.0, "small", seed=7)
    o = oracle_mod.Oracle(b)
    r = o.chol(tau=1e-6)
    V = dense_v(o, b, r.shell_pairs)
    return b, o, r, V


@pytest.mark.parametrize("flags", [0, 1])
def test_vectors_reproduce_pivot_columns_and_stay_within_tau(oracle_mod, system, flags):
    b, o, r0, V = system
    r = o.chol(tau=1e-6, flags=flags)
    assert r.status == 0 and r.nvecs > 0
    R = r.full_vectors()
    E = V - R.T @ R
    P = step2_columns(b, ref_meta(r))
    assert len(np.unique(P)) == r.nvecs
    vmax = np.abs(V).max()
    assert np.abs(E[:, P]).max() <= 1e-12 * vmax
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
    b, o, _, V = system
    r = o.chol(tau=1e-6, flags=1)
    P = step2_columns(b, ref_meta(r))
    assert np.array_equal(np.sort(P), np.sort(r.pivot_order - 1))


def test_subsets_do_not_change_the_decomposition(oracle_mod, system):
    b, o, r0, V = system
    r1 = o.chol(tau=1e-6, max_block_dim=200)
    assert r1.nsubsets > 1
    assert np.array_equal(r1.pivot_order, r0.pivot_order)
    R0, R1 = r0.full_vectors(), r1.full_vectors()
    assert np.abs(R0 - R1).max() <= 1e-12 * max(1.0, np.abs(R0).max())


def test_residual_bounds_pivot_count(oracle_mod, system):
    """A tighter threshold keeps more vectors and a smaller residual."""
    b, o, _, V = system
    ra, rb = o.chol(tau=1e-4), o.chol(tau=1e-8)
    assert ra.nvecs < rb.nvecs
    for r, tau in ((ra, 1e-4), (rb, 1e-8)):
        R = r.full_vectors()
        assert np.abs(np.diag(V - R.T @ R)).max() <= tau * (1 + 1e-9)
