"""Pins of two oracle functions by what the reference states or what the
mathematics fixes:

* or_pq2p_ge_q (chol2_pq2p_ge_q, TwoStepCholesky_Step2.f90:422-466) against
  the N = 3 traversal drawn in its own comment (:432-436),
      1
      2 5
      3 6 4
  (row p, column q), and, for every N up to 12, a bijection of
  1..N(N+1)/2 onto the lower triangle p >= q that differs from the Step-1
  storage order (b outer, a >= b inner) -- the mismatch behind the
  reference's Step-2 decode;
* or_mo_transform (chol_gammcor_MOTransf_2, Cholesky_Gammcor.f90:143-332)
  against the dense identity Rkab(k, a, b) = (CA^T L^k CB)(a, b), with L^k
  the symmetric AO matrix of vector k (L^k(p,q) = L^k(q,p) = R_k(pq)),
  including the columns the reference's pass bounds skip when a0 > 1
  (:190-197)."""
import numpy as np
import pytest

from gigpu import systems
from grad_helpers import tiny_basis


def test_pq2p_ge_q_drawn_traversal(oracle_mod):
    drawn = [[1, None, None], [2, 5, None], [3, 6, 4]]
    for p in range(1, 4):
        for q in range(1, p + 1):
            assert oracle_mod.pq2p_ge_q(drawn[p - 1][q - 1], 3) == (p, q)


@pytest.mark.parametrize("n", range(1, 13))
def test_pq2p_ge_q_is_a_bijection_onto_the_lower_triangle(oracle_mod, n):
    got = [oracle_mod.pq2p_ge_q(pq, n) for pq in range(1, n * (n + 1) // 2 + 1)]
    assert sorted(got) == sorted((p, q) for q in range(1, n + 1) for p in range(q, n + 1))
    step1 = [(p, q) for q in range(1, n + 1) for p in range(q, n + 1)]
    if n >= 3:
        assert got != step1   # Step 2 decodes a different order than Step 1 stores


def _ao_pairs(basis, shell_pairs):
    """AO indices (p, q) of every FULL orbital pair (b outer, a >= b inner for
    A == B), 0-based."""
    loc = basis.shell_loc() - 1   # ShellLoc is 1-based
    out = []
    for A, B in shell_pairs:
        na, nb = basis.shell_nang(A - 1), basis.shell_nang(B - 1)
        for b in range(nb):
            for a in range(b if A == B else 0, na):
                out.append((loc[A - 1] + a, loc[B - 1] + b))
    return np.array(out)


def _dense_rkab(basis, r, CA, a0, a1, CB, b0, b1, mirror=True):
    R = r.full_vectors()
    pq = _ao_pairs(basis, r.shell_pairs)
    n = basis.nao
    out = np.zeros((r.nvecs, a1 - a0 + 1, b1 - b0 + 1))
    A, B = CA[:, a0 - 1:a1], CB[:, b0 - 1:b1]
    for k in range(r.nvecs):
        Lk = np.zeros((n, n))
        Lk[pq[:, 0], pq[:, 1]] = R[k]
        if mirror:
            Lk[pq[:, 1], pq[:, 0]] = R[k]
        out[k] = A.T @ Lk @ B
    return out


@pytest.mark.parametrize("contracted", [False, True])
def test_mo_transform_equals_dense_identity(oracle_mod, contracted):
    b = tiny_basis(lmax=3, contracted=contracted) if contracted else systems.water_system(1, 6.0, "small", seed=3)
    o = oracle_mod.Oracle(b)
    r = o.chol(tau=1e-6)
    assert r.status == 0
    rng = np.random.default_rng(8)
    C, _ = np.linalg.qr(rng.normal(size=(b.nao, b.nao)))
    a0, a1, b0, b1 = 1, 5, 6, 14
    got = r.mo_transform(C, a0, a1, C, b0, b1).reshape(r.nvecs, a1 - a0 + 1, b1 - b0 + 1, order="F")
    ref = _dense_rkab(b, r, C, a0, a1, C, b0, b1)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    # planted: the (p != q) mirror term dropped (L^k one triangle only)
    wrong = _dense_rkab(b, r, C, a0, a1, C, b0, b1, mirror=False)
    assert np.abs(got - wrong).max() > 1e-3 * np.abs(ref).max()


def test_mo_transform_column_mask_of_the_pass_bounds(oracle_mod):
    """a0 > 1 with several passes: the reference's pass k covers
    a = (k-1) MaxNk + a0 .. min(k MaxNk, a1) (Cholesky_Gammcor.f90:190-197),
    so some a-columns are never transformed and stay zero (Rab = ZERO)."""
    b = systems.water_system(1, 6.0, "small", seed=3)
    o = oracle_mod.Oracle(b)
    r = o.chol(tau=1e-6)
    rng = np.random.default_rng(9)
    C, _ = np.linalg.qr(rng.normal(size=(b.nao, b.nao)))
    a0, a1, b0, b1 = 4, 15, 1, 6
    NA = a1 - a0 + 1
    maxnk = 3
    mb = maxnk * r.nvecs * b.nao * 8 / 2 ** 20
    mb_int = int(np.ceil(mb))
    maxnk = min(NA, (mb_int * 2 ** 20 // 8) // (r.nvecs * b.nao))
    got = r.mo_transform(C, a0, a1, C, b0, b1, max_buffer_mb=mb_int).reshape(r.nvecs, NA, b1 - b0 + 1, order="F")
    ref = _dense_rkab(b, r, C, a0, a1, C, b0, b1)
    npass = NA // maxnk + (NA % maxnk > 0)
    done = np.zeros(NA, bool)
    for k in range(1, npass + 1):
        lo, hi = (k - 1) * maxnk + a0, min(k * maxnk, a1)
        done[lo - a0:hi - a0 + 1] = True
    assert (~done).any() and done.any()
    assert np.abs(got[:, done] - ref[:, done]).max() <= 1e-13 * np.abs(ref).max()
    assert not got[:, ~done].any()
