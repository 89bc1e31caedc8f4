"""Tensor hypercontraction (THC) front end, mirroring THC_Gammcor.f90 and the
THC driver of TensorHypercontraction.f90 over the device library.

    thc_Chol_Pivots       TensorHypercontraction.f90:772-968: pruning of a parent grid
    thc_Grid              TensorHypercontraction.f90:300-390 from the caller's parent grid
    thc_CoulombMatrix     thc_CoulombMatrix_QuadraticMemory (TensorHypercontraction.f90:232-272)
                          with the grid points given by the caller
    thc_ReduceGrid        TensorHypercontraction.f90:275-297
    thc_gammcor_XZ        THC_Gammcor.f90:154-253 (accuracy presets, Omega -> Kappa)
    thc_gammcor_Rkab      THC_Gammcor.f90:25-80

The Becke/Lebedev parent grid (becke_MolecularGrid, called by thc_Grid) is
not part of the library: the caller gives the parent points (npts, 3) in
bohr.  Their pruning, the decomposition, the orbitals on the grid, the
least-squares Z and Rkab run on the device (gi_thc_grid_pivots,
gi_chol_build, gi_thc_orbitals, gi_thc_z, gi_thc_rkab).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import lib

THC_ACCURACY_DEFAULT = 1
THC_ACCURACY_TIGHT = 2
THC_ACCURACY_LUDICROUS = 3
THC_ACCURACY_DEBUG = 4

# (CholeskyTauThresh, QRThresh) per accuracy (THC_Gammcor.f90:196-212);
# QRThreshReduced = QRThresh there
_PRESETS = {THC_ACCURACY_DEFAULT: (1.0e-5, 1.0e-3), THC_ACCURACY_TIGHT: (1.0e-6, 1.0e-4),
            THC_ACCURACY_LUDICROUS: (1.0e-7, 1.0e-5), THC_ACCURACY_DEBUG: (1.0e-10, 1.0e-6)}


class THCError(RuntimeError):
    """error stop of the reference's THC routines."""


@dataclass
class TCoulTHCGrid:
    """TCoulTHCGrid: collocation Xgp (NGrid, NAO), Zgk (NGrid, NCholesky),
    ZgkReduced for the first NGridReduced points, Zgh = Zgk Zgk^T."""
    Xgp: np.ndarray
    Zgk: np.ndarray
    ZgkReduced: np.ndarray | None
    Zgh: np.ndarray | None
    NGrid: int
    NGridReduced: int


def thc_Chol_Pivots(ctx: "lib.Context", ParentXYZ, QRThresh: float, QRThreshReduced: float = -1.0,
                    BlockDim: int = 500):
    """(pivot points (NGrid, 3) in pivot order, NGridReduced) of the parent
    grid on the context's basis."""
    xyz = np.asarray(ParentXYZ, dtype=np.float64).reshape(-1, 3)
    try:
        piv, nred = ctx.thc_grid_pivots(xyz, QRThresh, QRThreshReduced, BlockDim)
    except lib.GigpuError as e:
        raise THCError(str(e))
    return xyz[piv], nred


def thc_Grid(ctx: "lib.Context", ParentXYZ, QRThresh: float = 1.0e-3, QRThreshReduced: float = 1.0e-3,
             BlockDim: int = 500):
    """(Xgp, points, NGrid, NGridReduced): the pruned grid and its
    row-normalized orbitals, thc_Grid after becke_MolecularGrid."""
    pts, nred = thc_Chol_Pivots(ctx, ParentXYZ, QRThresh, QRThreshReduced, BlockDim)
    n = len(pts)
    if QRThreshReduced <= 0.0:
        nred = n
    X = ctx.thc_orbitals(pts, normalize=True)
    return X, pts, n, nred


def thc_CoulombMatrix(AOBasis, GridXYZ, CholeskyTauThresh: float, Kappa: float = 0.0, NGridReduced: int = 0,
                      ctx: "lib.Context" = None, with_zgh: bool = True, MaxBlockDim: int = 4000,
                      CholVecsBlock: int = 500, QRThresh: float | None = None,
                      QRThreshReduced: float | None = None, BlockDim: int = 500) -> TCoulTHCGrid:
    """Cholesky decomposition, the grid (GridXYZ as given, or, with QRThresh,
    pruned from the parent grid GridXYZ as thc_Grid does, NGridReduced then
    from QRThreshReduced), its row-normalized orbitals and the least-squares
    Z; Zgh = Zgk Zgk^T when with_zgh."""
    own = ctx is None
    if own:
        ctx = lib.Context(0)
    try:
        if getattr(ctx, "basis", None) is not AOBasis:
            ctx.set_basis(AOBasis)
        ctx.chol_build(tau=CholeskyTauThresh, kappa=Kappa, max_block_dim=MaxBlockDim, vec_block=CholVecsBlock)
        xyz = np.asarray(GridXYZ, dtype=np.float64).reshape(-1, 3)
        try:
            if QRThresh is not None:
                red_thr = QRThresh if QRThreshReduced is None else QRThreshReduced
                X, xyz, n, NGridReduced = thc_Grid(ctx, xyz, QRThresh, red_thr, BlockDim)
            else:
                n = xyz.shape[0]
                X = ctx.thc_orbitals(xyz, normalize=True)
            Z, Zr = ctx.thc_z(X, NGridReduced)
        except lib.GigpuError as e:
            raise THCError(str(e))
        red = NGridReduced if 0 < NGridReduced < n else n
        return TCoulTHCGrid(Xgp=X, Zgk=Z, ZgkReduced=Zr, Zgh=(Z @ Z.T) if with_zgh else None, NGrid=n,
                            NGridReduced=red)
    finally:
        if own:
            ctx.close()


def thc_ReduceGrid(THCGrid: TCoulTHCGrid) -> None:
    """Keep only the first NGridReduced points: Zgk <- ZgkReduced, Xgp
    truncated, Zgh recomputed (in place)."""
    if THCGrid.NGridReduced < THCGrid.NGrid:
        n = THCGrid.NGridReduced
        THCGrid.NGrid = n
        THCGrid.Zgk, THCGrid.ZgkReduced = THCGrid.ZgkReduced, None
        THCGrid.Xgp = THCGrid.Xgp[:n].copy(order="F")
        THCGrid.Zgh = THCGrid.Zgk @ THCGrid.Zgk.T


def thc_gammcor_XZ(AOBasis, ParentXYZ, Accuracy: int, CholeskyThresh: float | None = None,
                   THCThresh: float | None = None, Omega: float | None = None, ctx: "lib.Context" = None):
    """(Xgp, Zgk) for the accuracy preset (THC_Gammcor.f90:154-253): the
    parent grid ParentXYZ pruned at QRThresh (THCThresh if given), Z on the
    pruned grid."""
    if Accuracy not in _PRESETS:
        raise THCError("Invalid THC accuracy value")
    tau = _PRESETS[Accuracy][0] if CholeskyThresh is None else CholeskyThresh
    qr = _PRESETS[Accuracy][1] if THCThresh is None else THCThresh
    kappa = 1.0 / Omega ** 2 if (Omega is not None and Omega > 0.0) else 0.0
    g = thc_CoulombMatrix(AOBasis, ParentXYZ, tau, kappa, ctx=ctx, with_zgh=False, QRThresh=qr, QRThreshReduced=qr)
    return g.Xgp, g.Zgk


def thc_gammcor_Rkab(CA, a0, a1, CB, b0, b1, AOBasis, ParentXYZ, Accuracy: int, ExternalOrdering: int,
                     ctx: "lib.Context" = None, Omega: float | None = None) -> np.ndarray:
    """Rkab (NCholesky, NA*NB), a fastest: sum_g Zgk(g,k) Xga(g,a) Xgb(g,b)
    with Xga = Xgp C_ao(:, a0:a1) on the pruned grid (THC_Gammcor.f90:25-80)."""
    own = ctx is None
    if own:
        ctx = lib.Context(0)
    try:
        X, Z = thc_gammcor_XZ(AOBasis, ParentXYZ, Accuracy, Omega=Omega, ctx=ctx)
        try:
            return ctx.thc_rkab(X, Z, CA, a0, a1, CB, b0, b1, ExternalOrdering)
        except lib.GigpuError as e:
            raise THCError(str(e))
    finally:
        if own:
            ctx.close()
