"""Pins of the THC oracle (oracle/thc.py) on the CPU.

orbitals: Gauss-Hermite quadrature of products of orbitals (one primitive
per shell, so the product times exp(p |r - P|^2) is a polynomial and the
quadrature is exact) reproduces the oracle's analytic overlap matrix, whose
own pin is Gauss-Hermite per dimension (test_oracle_onee.py); a contracted
s function at its centre is norm * sum(c).  z: the least-squares fit over
the full square of orbital pairs, solved by numpy's SVD lstsq.  rkab: the
oracle's MO transform of the fitted vectors."""
import math

import numpy as np
import pytest

import aomap
import pyoracle
import thc
from gigpu import systems
from gigpu.basis import build_basis


@pytest.fixture(scope="module", autouse=True)
def _built():
    pyoracle.build()


def gh_grid_3d(P, p, n=12):
    x, w = np.polynomial.hermite.hermgauss(n)
    g = np.stack(np.meshgrid(x, x, x, indexing="ij"), axis=-1).reshape(-1, 3)
    W = (w[:, None, None] * w[None, :, None] * w[None, None, :]).reshape(-1)
    return P + g / math.sqrt(p), W / p ** 1.5, g


@pytest.mark.parametrize("la,lb", [(0, 0), (1, 0), (1, 1), (2, 1), (2, 2), (3, 1), (3, 3)])
@pytest.mark.parametrize("spher", [True, False])
def test_orbitals_integrate_to_overlap(la, lb, spher):
    a, b = 0.7, 1.9
    coords = np.array([[0.1, -0.3, 0.2], [0.9, 0.5, -0.6]])
    basis = build_basis(["A", "B"], coords, {"A": [(la, [a], [1.0])], "B": [(lb, [b], [1.0])]}, spher_ao=spher)
    S, _, _ = pyoracle.Oracle(basis).one_electron(np.zeros(2))
    p = a + b
    P = (a * coords[0] + b * coords[1]) / p
    pts, W, g = gh_grid_3d(P, p)
    X = thc.orbitals(basis, pts, normalize=False)
    # X_i X_j exp(p |r - P|^2) is a polynomial times a constant: exact for
    # the cross block; the diagonal blocks use the same grid with exponents
    # 2a and 2b, so take only the A-B block here
    na = 2 * la + 1 if spher else (la + 1) * (la + 2) // 2
    w = W * np.exp((g * g).sum(axis=1))
    Sq = (X[:, :na] * w[:, None]).T @ X[:, na:]
    ref = S[:na, na:]
    assert np.abs(Sq - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-300)
    # one-centre blocks: grid around each centre with exponent 2a / 2b
    for lo, hi, e, C in ((0, na, a, coords[0]), (na, basis.nao, b, coords[1])):
        pts, W, g = gh_grid_3d(C, 2 * e)
        X = thc.orbitals(basis, pts, normalize=False)
        w = W * np.exp((g * g).sum(axis=1))
        Sq = (X[:, lo:hi] * w[:, None]).T @ X[:, lo:hi]
        assert np.abs(Sq - S[lo:hi, lo:hi]).max() <= 1e-12


def test_contracted_s_at_centre():
    ex, cf = [13.0, 2.1, 0.45], [0.15, 0.55, 0.48]
    coords = np.array([[0.2, 0.1, -0.4]])
    basis = build_basis(["A"], coords, {"A": [(0, ex, cf), (1, ex[:2], cf[:2])]})
    X = thc.orbitals(basis, coords, normalize=False)
    par = 0
    want = basis.norm_spher[par, 0] * basis.cntr_coeffs[par, :3].sum()
    assert abs(X[0, 0] - want) <= 1e-14 * abs(want)
    assert np.all(X[0, 1:] == 0.0)   # p functions vanish at their centre
    # and the function is a unit vector in the overlap metric
    S, _, _ = pyoracle.Oracle(basis).one_electron(np.zeros(1))
    assert abs(S[0, 0] - 1.0) < 1e-12


def test_normalized_rows():
    b = systems.water_system(1, 6.0, "small", seed=1)
    pts = systems.atom_centred_points(b, ndir=4, seed=2)
    X0 = thc.orbitals(b, pts, normalize=False)
    X = thc.orbitals(b, pts)
    assert np.allclose(np.linalg.norm(X, axis=1), 1.0, atol=1e-14)
    assert np.allclose(X * np.linalg.norm(X0, axis=1)[:, None], X0, rtol=1e-14, atol=1e-16)


@pytest.fixture(scope="module")
def water():
    b = systems.water_system(1, 6.0, "small", seed=1)
    r = pyoracle.Oracle(b).chol(tau=1e-6)
    assert r.status == 0
    cols = thc.storage_pairs(b, r.shell_pairs, r.shell_pair_loc, r.max_subset_dim, r.nsubsets)
    return b, r, cols


def full_square_r(basis, r, cols):
    """R(k, p, q) over all ordered AO pairs from the stored columns."""
    n = basis.nao
    R = np.zeros((r.nvecs, n, n))
    Rs = r.rkpq.reshape(-1, r.nvecs)
    for c, (p, q) in enumerate(cols):
        if p >= 0:
            R[:, p, q] = Rs[c]
            R[:, q, p] = Rs[c]
    return R


def test_storage_pairs_cover_every_pair_once(water):
    b, r, cols = water
    n = b.nao
    ok = cols[:, 0] >= 0
    key = np.maximum(cols[ok, 0], cols[ok, 1]) * n + np.minimum(cols[ok, 0], cols[ok, 1])
    assert len(np.unique(key)) == ok.sum() == n * (n + 1) // 2
    # the vectors reproduce the Step-2 pivot columns of (pq|rs) exactly: with
    # the (p, q) of the storage columns, sum_k R(k,p,q) R(k,pj,qj) = (pq|pj qj)
    from chol_helpers import dense_v, pair_rows
    o = pyoracle.Oracle(b)
    V = dense_v(o, b, r.shell_pairs)
    rows = pair_rows(b, r.shell_pairs)
    loc = b.shell_loc() - 1
    pq = np.array([(loc[r.shell_pairs[s][0] - 1] + a, loc[r.shell_pairs[s][1] - 1] + bb) for s, a, bb in rows])
    R = full_square_r(b, r, cols)
    P = pyoracle.step2_columns(b, {"shell_pairs": r.shell_pairs, "shell_pair_loc": r.shell_pair_loc,
                                   "pivot_shell_pairs": r.pivot_shell_pairs,
                                   "pivot_shell_pair_loc": r.pivot_shell_pair_loc,
                                   "pivot_shell_pair_dim": r.pivot_shell_pair_dim,
                                   "pivot_orb_pairs": r.pivot_orb_pairs})
    Rf = R[:, pq[:, 0], pq[:, 1]]
    G = Rf.T @ Rf[:, P]
    assert np.abs(G - V[:, P]).max() <= 1e-10 * np.abs(V).max()


def test_z_is_the_full_square_least_squares_fit(water):
    b, r, cols = water
    pts = systems.atom_centred_points(b, ndir=6, seed=3)
    X = thc.orbitals(b, pts)
    Z, _ = thc.z(X, cols, r.rkpq)
    n = b.nao
    A = np.einsum("gp,gq->pqg", X, X).reshape(n * n, -1)
    R = full_square_r(b, r, cols).reshape(r.nvecs, n * n).T
    Zl, *_ = np.linalg.lstsq(A, R, rcond=None)
    assert np.abs(Z - Zl).max() <= 1e-9 * np.abs(Zl).max()
    # a dropped 1/2 on the diagonal products moves the fit far beyond that
    bad = thc.xgpq(X, cols)
    bad[:, (cols[:, 0] == cols[:, 1]) & (cols[:, 0] >= 0)] *= 2.0
    XRb = 2.0 * bad @ r.rkpq.reshape(-1, r.nvecs)
    Zb = np.linalg.solve(thc.metric(X), XRb)
    assert np.abs(Zb - Zl).max() > 1e-3 * np.abs(Zl).max()


def test_reduced_grid(water):
    b, r, cols = water
    pts = systems.atom_centred_points(b, ndir=6, seed=3)
    X = thc.orbitals(b, pts)
    Z, Zr = thc.z(X, cols, r.rkpq, npts_reduced=40)
    Z40, _ = thc.z(X[:40], cols, r.rkpq)
    assert Zr.shape == (40, r.nvecs)
    assert np.abs(Zr - Z40).max() <= 1e-12 * np.abs(Z40).max()


@pytest.mark.parametrize("ordering", [aomap.AUTO2E, aomap.ORCA, aomap.MOLPRO])
def test_rkab_is_mo_transform_of_fitted_vectors(water, ordering):
    b, r, cols = water
    pts = systems.atom_centred_points(b, ndir=6, seed=3)
    X = thc.orbitals(b, pts)
    Z, _ = thc.z(X, cols, r.rkpq)
    rng = np.random.default_rng(7)
    CA = rng.standard_normal((b.nao, 9))
    CB = rng.standard_normal((b.nao, 7))
    CAo = thc.coefficients_to_auto2e(b, CA, ordering)
    CBo = thc.coefficients_to_auto2e(b, CB, ordering)
    got = thc.rkab(X, Z, CAo, 1, 8, CBo, 3, 7)
    # fitted vectors R~(k, pq) = sum_g Z(g,k) X(g,p) X(g,q) in the storage layout
    W = thc.xgpq(X, cols)
    W[:, (cols[:, 0] == cols[:, 1]) & (cols[:, 0] >= 0)] *= 2.0   # undo the halving
    fitted = (W.T @ Z).reshape(r.nsubsets, r.max_subset_dim, r.nvecs)
    r.set_rkpq(fitted)
    try:
        # a0 = 1: the reference's pass bounds of the MO transform skip columns
        # when a0 > 1 (Cholesky_Gammcor.f90:196-197); THC has no such passes
        ref = r.mo_transform(CAo, 1, 8, CBo, 3, 7)
    finally:
        r.set_rkpq(pyoracle.Oracle(b).chol(tau=1e-6).rkpq)
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


def dense_pivots(X, thresh, thresh_red):
    """Textbook pivoted Cholesky of the whole S = (X X^T)^2: no blocks, no
    candidate compression; first index of the largest remaining diagonal."""
    S = (X @ X.T) ** 2
    n = len(S)
    D = np.ones(n)
    cols, piv, nred = [], [], 0
    while len(piv) < n:
        h = int(np.argmax(D))
        if not D[h] > thresh:
            break
        c = S[:, h].copy()
        for v in cols:
            c -= v * v[h]
        c /= np.sqrt(D[h])
        piv.append(h)
        if D[h] > thresh_red:
            nred = len(piv)
        cols.append(c)
        D -= c * c
    return np.array(piv), nred


@pytest.fixture(scope="module")
def parent_grid():
    b = systems.water_system(2, 6.0, "small", seed=1)
    pts = systems.atom_centred_points(b, radii=(0.2, 0.5, 0.9, 1.4, 2.0, 3.0, 4.5), ndir=30, seed=5)
    return b, pts


@pytest.mark.parametrize("qr,block", [(1e-2, 100), (1e-3, 64), (1e-3, 5000)])
def test_grid_pivots_equal_dense_pivoted_cholesky(parent_grid, qr, block):
    b, pts = parent_grid
    piv, nred = thc.grid_pivots(b, pts, qr, 3 * qr, block_dim=block)
    X = thc.orbitals(b, pts)
    dpiv, dnred = dense_pivots(X, qr * qr, 9 * qr * qr)
    assert np.array_equal(piv, dpiv)
    assert nred == dnred and 0 < nred < len(piv)
    # every point left out has its Schur-complement diagonal below the threshold
    S = (X @ X.T) ** 2
    rest = np.setdiff1d(np.arange(len(pts)), piv)
    Spp = S[np.ix_(piv, piv)]
    Src = S[np.ix_(rest, piv)]
    schur = 1.0 - np.einsum("ij,ji->i", Src, np.linalg.solve(Spp, Src.T))
    assert schur.max() <= qr * qr * (1 + 1e-6)
    # an unsquared metric (the rank-revealing QR of X instead of X(g,pq))
    # picks another sequence
    ppiv, _ = dense_pivots_unsquared(X, qr * qr)
    assert not np.array_equal(ppiv[:len(piv)], piv)


def dense_pivots_unsquared(X, thresh):
    S = X @ X.T
    D = np.ones(len(S))
    cols, piv = [], []
    while len(piv) < len(S):
        h = int(np.argmax(D))
        if not D[h] > thresh:
            break
        c = S[:, h].copy()
        for v in cols:
            c -= v * v[h]
        c /= np.sqrt(D[h])
        piv.append(h)
        cols.append(c)
        D -= c * c
    return np.array(piv), len(piv)


def test_grid_pivots_without_reduced_threshold_and_last_block(parent_grid):
    b, pts = parent_grid
    piv, nred = thc.grid_pivots(b, pts, 1e-2, -1.0, block_dim=100)
    assert nred == len(piv)
    # six independent points, blocks of four: MaxIters = ceil(6 / 4) = 2 and
    # the second block may take only NGrid - NPivots = 2 more pivots; every
    # point is taken, in the order of the dense pivoted Cholesky
    six = b.atom_coords + np.array([0.3, -0.2, 0.25])
    p6, n6 = thc.grid_pivots(b, six, 1e-2, -1.0, block_dim=4)
    d6, _ = dense_pivots(thc.orbitals(b, six), 1e-4, -1.0)
    assert sorted(p6) == list(range(6)) and n6 == 6 and np.array_equal(p6, d6)
