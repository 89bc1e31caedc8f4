"""NumPy restatement of the tensor hypercontraction (THC) path.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product library.

    orbitals   gridfunc_Orbitals_Shell (GridFunctions.f90:313-390) for every
               shell (gridfunc_Orbitals, :154-240), then thc_normalize_Xgp
               (TensorHypercontraction.f90:600-637)
    xgpq       thc_Xgpq / thc_Xgpq_ab / thc_Xgpq_aa (TensorHypercontraction.f90:640-769)
    xr         thc_XR with the precomputed vectors (:479-577)
    metric     thc_S (:580-597)
    z          thc_Z (:393-476); the solve is LAPACK dsysv with uplo "L", the
               routine real_Axb_symmetric_sysv calls (real_linalg.f90:656-703)
    rkab       thc_gammcor_Xga + thc_gammcor_Rkab_2 (THC_Gammcor.f90:82-106, 175-191)
    grid_pivots  thc_Chol_Pivots + thc_Chol_NextBlock + thc_Chol_Compress
               (TensorHypercontraction.f90:772-1127): pivoted Cholesky of S built
               column by column, candidates compressed after every block

Pins (tests/test_oracle_thc.py): the orbitals integrate to the oracle's
analytic overlap matrix under Gauss-Hermite quadrature (exact for one
primitive per shell, every l, spherical and Cartesian); a contracted s
function at its centre equals norm * sum(c); Z equals the least-squares fit
of R(k, pq) by sum_g X(g,p) X(g,q) Z(g,k) over the FULL square of orbital
pairs (numpy lstsq, no symmetry folding, no factor 2, no halved diagonal);
Rkab equals the oracle's MO transform of the fitted vectors; grid_pivots
gives the pivot sequence of a dense textbook pivoted Cholesky of the whole
S (no blocks, no compression) and leaves every Schur-complement diagonal
below the threshold.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg.lapack as lapack

import aomap
import pyoracle


def cart_exponents(l):
    """(lx, ly, lz) of the Cartesian functions of a shell in Auto2e order:
    lx = l..0, then ly = l - lx..0 (Auto2eGlobal.py:35-43)."""
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def orbitals(basis, xyz, normalize=True):
    """Xgp (npts, nao) of the points xyz (npts, 3), Auto2e AO order."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    X = np.zeros((xyz.shape[0], basis.nao))
    loc = basis.shell_loc()
    for s in range(basis.nshells):
        par = int(basis.shell_params_idx[s]) - 1
        l = int(basis.shell_momentum[par])
        npr = int(basis.nprimitives[par])
        A = basis.atom_coords[int(basis.shell_centers[s]) - 1]
        d = xyz - A
        r2 = (d * d).sum(axis=1)
        rad = np.zeros(len(r2))
        for i in range(npr):
            rad += basis.cntr_coeffs[par, i] * np.exp(-basis.exponents[par, i] * r2)
        cart = np.stack([d[:, 0] ** lx * d[:, 1] ** ly * d[:, 2] ** lz * rad
                         for lx, ly, lz in cart_exponents(l)], axis=1)
        p0 = int(loc[s]) - 1
        if basis.spher_ao:
            vals = cart @ pyoracle.spher_u(l).T if l >= 2 else cart
            nrm = basis.norm_spher[par, :2 * l + 1]
        else:
            vals = cart
            nrm = basis.norm_cart[par, :cart.shape[1]]
        X[:, p0:p0 + vals.shape[1]] = vals * nrm
    if normalize:
        n = np.sqrt((X * X).sum(axis=1))
        X = X * (1.0 / n)[:, None]
    return X


def storage_pairs(basis, shell_pairs, shell_pair_loc, max_subset_dim, nsubsets):
    """(p, q) (0-based AOs) of every stored column (X - 1) msd + position of
    the Cholesky vectors, -1 for padding: for shell pair (A, B), b outer,
    a = (A == B ? b : 0).. inner."""
    cols = -np.ones((max_subset_dim * nsubsets, 2), dtype=np.int64)
    loc = basis.shell_loc()
    for s in range(len(shell_pairs)):
        A, B = (int(v) - 1 for v in shell_pairs[s])
        pos0, X = int(shell_pair_loc[s][1]) - 1, int(shell_pair_loc[s][2])
        na, nb = basis.shell_nang(A), basis.shell_nang(B)
        qq = 0
        for b in range(nb):
            for a in range(b if A == B else 0, na):
                cols[(X - 1) * max_subset_dim + pos0 + qq] = (loc[A] - 1 + a, loc[B] - 1 + b)
                qq += 1
    return cols


def xgpq(X, cols):
    """Xgpq (npts, ncols): X(g,p) X(g,q), halved for p = q, 0 for padding."""
    p, q = cols[:, 0], cols[:, 1]
    ok = p >= 0
    W = np.zeros((X.shape[0], len(cols)))
    W[:, ok] = X[:, p[ok]] * X[:, q[ok]]
    W[:, ok & (p == q)] *= 0.5
    return W


def xr(X, cols, rkpq):
    """XR(g, k) = 2 sum_c Xgpq(g, c) R(k, c); rkpq[X, i, k] = Rkpq(k, i+1, X+1)."""
    R = rkpq.reshape(-1, rkpq.shape[2])          # (ncols, nvecs), column c = X msd + i
    return 2.0 * xgpq(X, cols) @ R


def metric(X, n=None):
    Y = X if n is None else X[:n]
    return (Y @ Y.T) ** 2


def _sysv_lower(S, B):
    lu, piv, x, info = lapack.dsysv(S, B, lower=1)
    if info != 0:
        raise RuntimeError("Linear system solver returned with info=%d" % info)
    return x


def z(X, cols, rkpq, npts_reduced=0):
    """(Zgk, Zgk_reduced or None) as thc_Z computes them."""
    XR = xr(X, cols, rkpq)
    Zr = None
    n = X.shape[0]
    if 0 < npts_reduced < n:
        Zr = _sysv_lower(metric(X, npts_reduced), XR[:npts_reduced].copy())
    return _sysv_lower(metric(X), XR), Zr


def rkab(X, Zgk, CA_ao, a0, a1, CB_ao, b0, b1):
    """Rkab (nvecs, NA*NB), a fastest, from AO-order coefficients."""
    Xa = X @ CA_ao[:, a0 - 1:a1]
    Xb = X @ CB_ao[:, b0 - 1:b1]
    W = (Xb[:, :, None] * Xa[:, None, :]).reshape(X.shape[0], -1)   # column a + NA b
    return Zgk.T @ W


def coefficients_to_auto2e(basis, C_ext, ordering):
    """auto2e_interface_C: rows from the external AO order to Auto2e order,
    Orca signs applied (the oracle's AO map)."""
    if ordering == aomap.AUTO2E:
        return np.array(C_ext, dtype=np.float64)
    shell_l, shell_atom = aomap.basis_shells(basis)
    m, ph = aomap.ao_map(shell_l, shell_atom, basis.spher_ao, ordering)
    out = np.zeros_like(np.asarray(C_ext, dtype=np.float64))
    out[np.asarray(m) - 1] = np.asarray(C_ext)
    return out * np.asarray(ph)[:, None]


def grid_pivots(basis, xyz, qr_thresh, qr_thresh_reduced=-1.0, block_dim=500):
    """(pivots, npivots_reduced): 0-based indices into xyz of the pruned grid
    in pivot order, and NGridReduced, as thc_Chol_Pivots computes them.
    Raises when MaxIters blocks do not converge."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    ngrid = xyz.shape[0]
    thresh = qr_thresh ** 2
    thresh_red = qr_thresh_reduced ** 2 if qr_thresh_reduced > 0 else -1.0
    cand = np.arange(ngrid)
    D = np.ones(ngrid)                 # normalized rows: S(g, g) = 1
    L = np.zeros((ngrid, 0))           # vectors over the current candidates
    piv, nred = [], 0
    max_iters = ngrid // block_dim + (ngrid % block_dim > 0)
    converged = False
    for _ in range(max_iters):
        Xc = orbitals(basis, xyz[cand])
        max_j = min(block_dim, ngrid - len(piv))
        cols = []
        for _k in range(max_j):
            h = int(np.argmax(D))      # first index of the maximum, as maxloc
            dh = D[h]
            if not dh > thresh:
                break
            piv.append(int(cand[h]))
            if dh > thresh_red:
                nred = len(piv)
            lj = (Xc @ Xc[h]) ** 2
            if L.shape[1]:
                lj -= L @ L[h]
            if cols:
                Lc = np.stack(cols, axis=1)
                lj -= Lc @ Lc[h]
            lj /= np.sqrt(dh)
            cols.append(lj)
            D = D - lj * lj
        if cols:
            L = np.concatenate([L, np.stack(cols, axis=1)], axis=1)
        keep = D > thresh
        cand, D, L = cand[keep], D[keep], L[keep]
        if len(cand) == 0:
            converged = True
            break
    if not converged:
        raise RuntimeError("Cholesky decomposistion failed to converge")
    if qr_thresh_reduced <= 0:
        nred = len(piv)
    return np.array(piv, dtype=np.int64), nred
