"""THC on the GPU (gi_thc_orbitals, gi_thc_z, gi_thc_rkab) against the oracle
(oracle/thc.py), and at the 1200-AO size through properties: orbitals
point by point, the normal equations S Z = XR with XR rows from the
independent MO-transform path, Rkab from the device X and Z."""
import numpy as np
import pytest

import aomap
import pyoracle
import thc
from gigpu import lib, systems
from gigpu import thc as gthc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    pyoracle.build()
    c = lib.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("basis,spher", [("large", True), ("large", False), ("gh", True)])
def test_orbitals_vs_oracle(ctx, basis, spher):
    b = systems.water_system(2, 6.0, basis, seed=11, spher_ao=spher, contracted=True)
    ctx.set_basis(b)
    pts = systems.atom_centred_points(b, ndir=10, seed=5)
    print("orbitals", basis, spher, b.nao, len(pts), flush=True)
    for norm in (False, True):
        X = ctx.thc_orbitals(pts, normalize=norm)
        Xo = thc.orbitals(b, pts, normalize=norm)
        assert np.abs(X - Xo).max() <= 1e-13 * np.abs(Xo).max()


def test_orbitals_zero_row_fails(ctx):
    b = systems.water_system(1, 6.0, "small", seed=1)
    ctx.set_basis(b)
    far = np.array([[0.0, 0.0, 0.0], [1e4, 0.0, 0.0]])
    with pytest.raises(lib.GigpuError, match="cannot be normalized"):
        ctx.thc_orbitals(far, normalize=True)
    X = ctx.thc_orbitals(far, normalize=False)
    assert np.all(X[1] == 0.0)


def _oracle_z(ctx, b, X, nred=0):
    m = ctx.chol_meta()
    cols = thc.storage_pairs(b, m["shell_pairs"], m["shell_pair_loc"], m["max_subset_dim"], m["nsubsets"])
    return thc.z(X, cols, ctx.chol_rkpq(), nred)


@pytest.mark.parametrize("solver", ["cholesky", "lu"])
@pytest.mark.parametrize("ndir,tol", [(6, 1e-9), (12, 1e-5)])
def test_z_vs_oracle(ctx, ndir, tol, solver):
    """Z from the device's own decomposition against the oracle's dsysv
    solve of the same system, by the library's Cholesky of S and by its LU
    (option thc_cholesky = 0): the solvers agree to about cond(S) * 1e-16
    (cond(S) ~ 1e5 for 72 points, ~ 1e9 for 144)."""
    b = systems.water_system(1, 6.0, "small", seed=1)
    ctx.set_basis(b)
    ctx.chol_build(tau=1e-6)
    pts = systems.atom_centred_points(b, ndir=ndir, seed=3)
    print("z", ndir, len(pts), solver, flush=True)
    X = ctx.thc_orbitals(pts)
    ctx.set_option("thc_cholesky", 1 if solver == "cholesky" else 0)
    try:
        Z, Zr = ctx.thc_z(X, npts_reduced=len(pts) // 2)
        # npts_reduced >= npts: no reduced system, the same Z
        Z2, Zr2 = ctx.thc_z(X, npts_reduced=len(pts))
    finally:
        ctx.set_option("thc_cholesky", 1)
    Zo, Zro = _oracle_z(ctx, b, X, len(pts) // 2)
    assert np.abs(Z - Zo).max() <= tol * np.abs(Zo).max()
    assert np.abs(Zr - Zro).max() <= tol * np.abs(Zro).max()
    assert Zr2 is None and np.array_equal(Z2, Z)


def test_z_contracted_with_range_separation(ctx):
    b = systems.water_system(2, 6.0, "small", seed=4, contracted=True)
    ctx.set_basis(b)
    ctx.chol_build(tau=1e-7, kappa=1.0 / 0.4 ** 2)
    pts = systems.atom_centred_points(b, ndir=5, seed=8)
    X = ctx.thc_orbitals(pts)
    Z, _ = ctx.thc_z(X)
    Zo, _ = _oracle_z(ctx, b, X)
    assert np.abs(Z - Zo).max() <= 1e-8 * np.abs(Zo).max()


@pytest.mark.parametrize("ordering", [aomap.AUTO2E, aomap.ORCA, aomap.PYSCF])
def test_rkab_vs_oracle(ctx, ordering):
    b = systems.water_system(2, 6.0, "large", seed=12)
    ctx.set_basis(b)
    ctx.chol_build(tau=1e-5)
    pts = systems.atom_centred_points(b, ndir=6, seed=9)
    X = ctx.thc_orbitals(pts)
    Z, _ = ctx.thc_z(X)
    rng = np.random.default_rng(ordering)
    CA = rng.standard_normal((b.nao, 12))
    CB = rng.standard_normal((b.nao, 9))
    got = ctx.thc_rkab(X, Z, CA, 3, 12, CB, 2, 9, ordering)
    ref = thc.rkab(X, Z, thc.coefficients_to_auto2e(b, CA, ordering), 3, 12,
                   thc.coefficients_to_auto2e(b, CB, ordering), 2, 9)
    assert got.shape == (Z.shape[1], 10 * 8)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_python_front_end(ctx):
    b = systems.water_system(1, 6.0, "small", seed=1)
    pts = systems.atom_centred_points(b, ndir=6, seed=3)
    g = gthc.thc_CoulombMatrix(b, pts, 1e-6, NGridReduced=30, ctx=ctx)
    assert g.NGrid == len(pts) and g.NGridReduced == 30 and g.ZgkReduced.shape[0] == 30
    assert np.allclose(g.Zgh, g.Zgk @ g.Zgk.T, rtol=0, atol=1e-12 * np.abs(g.Zgh).max())
    Zr = g.ZgkReduced.copy()
    gthc.thc_ReduceGrid(g)
    assert g.NGrid == 30 and np.array_equal(g.Zgk, Zr) and g.Xgp.shape == (30, b.nao)
    assert np.allclose(g.Zgh, Zr @ Zr.T, rtol=0, atol=1e-12 * np.abs(g.Zgh).max())
    CA = np.random.default_rng(2).standard_normal((b.nao, 6))
    R = gthc.thc_gammcor_Rkab(CA, 1, 6, CA, 1, 6, b, pts, gthc.THC_ACCURACY_TIGHT, lib.ORDERING_ORCA, ctx=ctx)
    assert R.shape[1] == 36
    # symmetric in (a, b) for CA = CB
    nv = R.shape[0]
    R3 = R.reshape(nv, 6, 6)
    assert np.abs(R3 - R3.transpose(0, 2, 1)).max() <= 1e-12 * np.abs(R3).max()
    with pytest.raises(gthc.THCError):
        gthc.thc_gammcor_XZ(b, pts, 9, ctx=ctx)


def test_z_requires_decomposition():
    b = systems.water_system(1, 6.0, "small", seed=1)
    c = lib.Context(0)
    try:
        c.set_basis(b)
        X = c.thc_orbitals(systems.atom_centred_points(b, ndir=4, seed=1))
        with pytest.raises(lib.GigpuError, match="gi_chol_build"):
            c.thc_z(X)
    finally:
        c.close()


def test_full_size_normal_equations(ctx):
    """(H2O)16, O 8s5p3d1f / H 4s2p1d, 1200 AOs, 48 atoms x 4 radii x 16
    directions = 3072 points: orbitals against the oracle, S Z = XR with XR
    rows from gi_mo_transform (C = rows of X: Rkab(k, g, g) = x_g^T R_k x_g),
    Rkab against the product of the device X and Z."""
    b = systems.config_large_cluster()
    ctx.set_basis(b)
    ctx.chol_build(tau=1e-5)
    pts = systems.atom_centred_points(b, ndir=16, seed=21)
    print("full size", b.nao, len(pts), ctx.chol_dims()["nvecs"], flush=True)
    X = ctx.thc_orbitals(pts)
    Xo = thc.orbitals(b, pts)
    assert np.abs(X - Xo).max() <= 1e-13
    Z, _ = ctx.thc_z(X)
    sel = np.array([0, 5, 700, 1500, 2222, 3071])
    Cg = np.asfortranarray(X[sel].T)
    Rk = ctx.mo_transform(Cg, 1, len(sel), Cg, 1, len(sel))
    nv = Rk.shape[0]
    XR = np.stack([Rk.reshape(nv, len(sel), len(sel))[:, i, i] for i in range(len(sel))])
    Y = X[sel] @ X.T
    SZ = (Y * Y) @ Z
    # a backward-stable solve leaves a residual of a few eps |S| |Z| (measured
    # 2.4e-13 against max|XR| = 1.4, max|Z| = 48 for this grid)
    err, zmax = np.abs(SZ - XR).max(), np.abs(Z).max()
    print("residual %.3e  max|XR| %.3e  max|Z| %.3e" % (err, np.abs(XR).max(), zmax), flush=True)
    assert err <= 1e-10 * np.abs(XR).max()
    rng = np.random.default_rng(4)
    CA = rng.standard_normal((b.nao, 8))
    got = ctx.thc_rkab(X, Z, CA, 1, 8, CA, 4, 7, 0)
    Xa = X @ CA
    ref = Z.T @ (Xa[:, 3:7, None] * Xa[:, None, :]).reshape(len(pts), -1)
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("schur_mb", [24576, 0, 5])
@pytest.mark.parametrize("qr,block", [(1e-2, 100), (1e-3, 64), (1e-3, 5000)])
def test_grid_pivots_vs_oracle(ctx, qr, block, schur_mb):
    """The pruned grid (pivot order and NGridReduced) equals the oracle's
    restatement of thc_Chol_Pivots on a 1260-point parent grid, with the
    explicit Schur complement from the start (the default budget), with
    on-the-fly columns throughout (budget 0) and with a switch from the
    one to the other once fewer than 809 candidates remain (5 MiB)."""
    b = systems.water_system(2, 6.0, "small", seed=1)
    ctx.set_basis(b)
    ctx.set_option("thc_schur_mb", schur_mb)
    pts = systems.atom_centred_points(b, radii=(0.2, 0.5, 0.9, 1.4, 2.0, 3.0, 4.5), ndir=30, seed=5)
    print("pivots", qr, block, schur_mb, len(pts), flush=True)
    try:
        piv, nred = ctx.thc_grid_pivots(pts, qr, 3 * qr, block)
        pv2, nred2 = ctx.thc_grid_pivots(pts, qr, -1.0, block)
    finally:
        ctx.set_option("thc_schur_mb", 24576)
    opiv, onred = thc.grid_pivots(b, pts, qr, 3 * qr, block)
    assert np.array_equal(piv, opiv)
    assert nred == onred
    assert np.array_equal(pv2, piv) and nred2 == len(piv)


def test_thc_grid_front_end(ctx):
    b = systems.water_system(1, 6.0, "small", seed=1)
    ctx.set_basis(b)
    parent = systems.atom_centred_points(b, radii=(0.3, 0.8, 1.5, 2.5, 4.0), ndir=24, seed=6)
    X, pts, n, nred = gthc.thc_Grid(ctx, parent, 1e-3, 3e-3, BlockDim=50)
    assert X.shape == (n, b.nao) and pts.shape == (n, 3) and 0 < nred <= n < len(parent)
    assert np.allclose(np.linalg.norm(X, axis=1), 1.0, atol=1e-14)
    ctx.chol_build(tau=1e-6)
    Z, Zr = ctx.thc_z(X, npts_reduced=nred)
    Zo, Zro = _oracle_z(ctx, b, X, nred)
    assert np.abs(Z - Zo).max() <= 1e-6 * np.abs(Zo).max()
    with pytest.raises(gthc.THCError):
        gthc.thc_Chol_Pivots(ctx, parent, -1.0)


def test_cartesian_z_rkab_and_pivots(ctx):
    """Cartesian contracted AOs (Molpro ordering for Rkab, the only external
    order the reference defines for Cartesian functions): pruning, Z and
    Rkab against the oracle."""
    b = systems.water_system(1, 6.0, "small", seed=7, spher_ao=False, contracted=True)
    ctx.set_basis(b)
    parent = systems.atom_centred_points(b, radii=(0.3, 0.8, 1.5, 2.5, 4.0), ndir=20, seed=4)
    piv, nred = ctx.thc_grid_pivots(parent, 1e-2, -1.0, 40)
    opiv, _ = thc.grid_pivots(b, parent, 1e-2, -1.0, 40)
    assert np.array_equal(piv, opiv) and nred == len(piv)
    pts = parent[piv][:60]
    # Cartesian d shells make the reference's Step-2 decode meet a singular
    # pivot block; the Step-1 decode factorizes the Step-1 pivots
    ctx.chol_build(tau=1e-6, flags=lib.CHOL_DECODE_FIXED)
    X = ctx.thc_orbitals(pts)
    assert np.abs(X - thc.orbitals(b, pts)).max() <= 1e-13
    Z, _ = ctx.thc_z(X)
    Zo, _ = _oracle_z(ctx, b, X)
    assert np.abs(Z - Zo).max() <= 1e-7 * np.abs(Zo).max()
    C = np.random.default_rng(5).standard_normal((b.nao, 5))
    got = ctx.thc_rkab(X, Z, C, 1, 5, C, 2, 4, aomap.MOLPRO)
    Co = thc.coefficients_to_auto2e(b, C, aomap.MOLPRO)
    ref = thc.rkab(X, Z, Co, 1, 5, Co, 2, 4)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_thc_error_paths(ctx):
    """Argument errors come back as statuses with the reference's messages
    where it has one, and nothing is written."""
    import ctypes as C
    b = systems.water_system(1, 6.0, "small", seed=1)
    ctx.set_basis(b)
    ctx.chol_build(tau=1e-6)
    pts = systems.atom_centred_points(b, ndir=4, seed=1)
    X = np.asfortranarray(ctx.thc_orbitals(pts))
    n, nv = X.shape[0], ctx.chol_dims()["nvecs"]
    L = ctx.L
    Z = np.zeros((n, nv), order="F")
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    # leading dimensions smaller than the number of points
    assert L.gi_thc_z(ctx.h, n, dp(X), n - 1, 0, dp(Z), n, None, 1) == 81
    assert L.gi_thc_z(ctx.h, n, dp(X), n, 0, dp(Z), n - 1, None, 1) == 81
    assert L.gi_thc_z(ctx.h, 0, dp(X), n, 0, dp(Z), n, None, 1) == 81
    assert not Z.any()
    # Rkab: empty column ranges, short CA
    CA = np.zeros((b.nao, 4), order="F")
    out = np.zeros(nv * 16)
    Zr, _ = ctx.thc_z(X)
    Zr = np.asfortranarray(Zr)
    assert L.gi_thc_rkab(ctx.h, n, dp(X), n, dp(Zr), n, nv, dp(CA), b.nao, 3, 2, dp(CA), b.nao, 1, 4, 0,
                         dp(out)) == 51
    assert L.gi_thc_rkab(ctx.h, n, dp(X), n, dp(Zr), n, nv, dp(CA), b.nao - 1, 1, 4, dp(CA), b.nao, 1, 4, 0,
                         dp(out)) == 52
    # pruning: thresholds and block size
    with pytest.raises(lib.GigpuError, match="QR threshold"):
        ctx.thc_grid_pivots(pts, 0.0)
    with pytest.raises(lib.GigpuError, match="block dimension"):
        ctx.thc_grid_pivots(pts, 1e-3, -1.0, 0)
    with pytest.raises(lib.GigpuError, match="Invalid external orbital ordering|ordering"):
        ctx.thc_rkab(X, Zr, CA, 1, 4, CA, 1, 4, 9)
