"""Dense-matrix helpers for checking two-step Cholesky results.

Orbital pairs are indexed by their FULL storage location (1-based in the
reference, 0-based here): shell pairs in ShellPairs order, inside a shell
pair a fastest, same-shell pairs as the column-major lower triangle
(TwoStepCholesky_Step1.f90, chol2_write_V*)."""
import numpy as np

from pyoracle import step2_columns  # noqa: F401  (re-exported for the tests)


def pair_rows(basis, shell_pairs):
    """[(s, a, b)] in FULL order; shell_pairs is ShellPairs (1-based shells)."""
    rows = []
    for s, (A, B) in enumerate(shell_pairs):
        na, nb = basis.shell_nang(A - 1), basis.shell_nang(B - 1)
        for bb in range(nb):
            for a in range(bb if A == B else 0, na):
                rows.append((s, a, bb))
    return rows


def dense_v(o, basis, shell_pairs):
    """(pq|rs) over all BASE orbital pairs, FULL order, from the oracle's ERIs."""
    rows = pair_rows(basis, shell_pairs)
    n = len(rows)
    V = np.zeros((n, n))
    blocks = {}
    for s, (A, B) in enumerate(shell_pairs):
        for t, (C_, D) in enumerate(shell_pairs):
            if t > s:
                continue
            blocks[(s, t)] = o.eri(A - 1, B - 1, C_ - 1, D - 1)
    for i, (s, a, b) in enumerate(rows):
        for j, (t, c, d) in enumerate(rows):
            V[i, j] = blocks[(s, t)][a, b, c, d] if t <= s else blocks[(t, s)][c, d, a, b]
    return V


def cholesky_factor(inv_l):
    """L with L L^T = V(P, P) from Inv_L = L^-1 (lower triangular)."""
    n = inv_l.shape[0]
    return np.linalg.solve(np.tril(inv_l), np.eye(n))
