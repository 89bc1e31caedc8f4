"""Tensor hypercontraction (THC) front end, mirroring THC_Gammcor.f90 and the
THC driver of TensorHypercontraction.f90 over the device library.

    thc_CoulombMatrix     thc_CoulombMatrix_QuadraticMemory (TensorHypercontraction.f90:232-272)
                          with the grid points given by the caller
    thc_gammcor_XZ        THC_Gammcor.f90:154-253 (accuracy presets, Omega -> Kappa)
    thc_gammcor_Rkab      THC_Gammcor.f90:25-80

The Becke/Lebedev parent grid and its Cholesky pruning (thc_Grid,
thc_Chol_Pivots, TensorHypercontraction.f90:300-390, 772-968) are not part
of the library: every call takes the grid points (npts, 3) in bohr.  The
decomposition, the orbitals on the grid, the least-squares Z and Rkab run on
the device (gi_chol_build, gi_thc_orbitals, gi_thc_z, gi_thc_rkab).
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
# QRThresh drives the grid pruning, which the caller does
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


def thc_CoulombMatrix(AOBasis, GridXYZ, CholeskyTauThresh: float, Kappa: float = 0.0, NGridReduced: int = 0,
                      ctx: "lib.Context" = None, with_zgh: bool = True, MaxBlockDim: int = 4000,
                      CholVecsBlock: int = 500) -> TCoulTHCGrid:
    """Cholesky decomposition, orbitals on the grid (row-normalized) and the
    least-squares Z; Zgh = Zgk Zgk^T when with_zgh."""
    own = ctx is None
    if own:
        ctx = lib.Context(0)
    try:
        if getattr(ctx, "basis", None) is not AOBasis:
            ctx.set_basis(AOBasis)
        ctx.chol_build(tau=CholeskyTauThresh, kappa=Kappa, max_block_dim=MaxBlockDim, vec_block=CholVecsBlock)
        xyz = np.asarray(GridXYZ, dtype=np.float64).reshape(-1, 3)
        n = xyz.shape[0]
        try:
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


def thc_gammcor_XZ(AOBasis, GridXYZ, Accuracy: int, CholeskyThresh: float | None = None,
                   Omega: float | None = None, ctx: "lib.Context" = None):
    """(Xgp, Zgk) for the accuracy preset (THC_Gammcor.f90:154-253)."""
    if Accuracy not in _PRESETS:
        raise THCError("Invalid THC accuracy value")
    tau = _PRESETS[Accuracy][0] if CholeskyThresh is None else CholeskyThresh
    kappa = 1.0 / Omega ** 2 if (Omega is not None and Omega > 0.0) else 0.0
    g = thc_CoulombMatrix(AOBasis, GridXYZ, tau, kappa, ctx=ctx, with_zgh=False)
    return g.Xgp, g.Zgk


def thc_gammcor_Rkab(CA, a0, a1, CB, b0, b1, AOBasis, GridXYZ, Accuracy: int, ExternalOrdering: int,
                     ctx: "lib.Context" = None, Omega: float | None = None) -> np.ndarray:
    """Rkab (NCholesky, NA*NB), a fastest: sum_g Zgk(g,k) Xga(g,a) Xgb(g,b)
    with Xga = Xgp C_ao(:, a0:a1) (THC_Gammcor.f90:25-80)."""
    own = ctx is None
    if own:
        ctx = lib.Context(0)
    try:
        X, Z = thc_gammcor_XZ(AOBasis, GridXYZ, Accuracy, Omega=Omega, ctx=ctx)
        try:
            return ctx.thc_rkab(X, Z, CA, a0, a1, CB, b0, b1, ExternalOrdering)
        except lib.GigpuError as e:
            raise THCError(str(e))
    finally:
        if own:
            ctx.close()
