"""Host-side mirror of the reference's ``Cholesky_Gammcor`` module
(src/gammcor-interface/Cholesky_Gammcor.f90), running on libgigpu.

Same names, argument meaning and error behaviour as the Fortran module:

* ``chol_gammcor_Rkpq(AOBasis, Accuracy, Omega=None)`` -> ``TCholeskyVecsOTF``
  (Cholesky_Gammcor.f90:335-377, :380-441)
* ``chol_gammcor_Rkab(Rkab, CA, a0, a1, CB, b0, b1, MaxBufferDimMB,
  CholeskyVecsOTF, AOBasis, ExternalOrdering)`` (Cholesky_Gammcor.f90:26-92)

Errors the reference reports with ``msg(..., MSG_ERROR); error stop`` raise
``GammcorError`` here with the same message text.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import lib

CHOL_ACCURACY_DEFAULT = 1
CHOL_ACCURACY_TIGHT = 2
CHOL_ACCURACY_LUDICROUS = 3
CHOL_ACCURACY_DEBUG = 4

# Auto2eInterface.f90:12-17
ORBITAL_ORDERING_AUTO2E = 0
ORBITAL_ORDERING_MOLPRO = 1
ORBITAL_ORDERING_ORCA = 2
ORBITAL_ORDERING_DALTON = 3
ORBITAL_ORDERING_OPEN_MOLCAS = 4
ORBITAL_ORDERING_PYSCF = 5


class GammcorError(RuntimeError):
    pass


@dataclass
class TChol2Params:
    """TwoStepCholesky_definitions.f90:20-53."""
    CholeskyTauThresh: float = 1.0e-7
    Kappa: float = 0.0
    MaxBlockDim: int = 4000
    CholVecsBlock: int = 500


@dataclass
class TChol2Vecs:
    """TwoStepCholesky_definitions.f90:56-74 (arrays hold the reference's
    1-based values; Inv_L stays on the device and is fetched on demand)."""
    NVecs: int = 0
    MaxSubsetDim: int = 0
    NSubsets: tuple = (0, 1)
    NOrbPairs: int = 0
    NShellPairs: int = 0
    ShellPairs: np.ndarray = field(default=None)       # (NShellPairs, 2)
    ShellPairLoc: np.ndarray = field(default=None)     # (NShellPairs, 3)
    ShellPairDim: np.ndarray = field(default=None)
    SubsetDim: np.ndarray = field(default=None)
    SubsetBounds: np.ndarray = field(default=None)     # (NSubsets, 2)
    Pivots: np.ndarray = field(default=None)
    PivotShellPairs: np.ndarray = field(default=None)
    PivotShellPairLoc: np.ndarray = field(default=None)
    PivotShellPairDim: np.ndarray = field(default=None)
    PivotOrbPairs: np.ndarray = field(default=None)
    NPivotShellPairs: int = 0
    pivot_order: np.ndarray = field(default=None)


@dataclass
class TCholeskyVecsOTF:
    """Cholesky_Gammcor.f90:19-22.  ``Rkpq`` (NVecs, MaxSubsetDim, NSubsets)
    is kept on the device; ``rkpq()`` copies it to the host as a
    (NSubsets, MaxSubsetDim, NVecs) C-order array (same memory layout)."""
    Chol2Data: TChol2Vecs
    ctx: "lib.Context"
    timings: dict = field(default_factory=dict)

    def rkpq(self) -> np.ndarray:
        return self.ctx.chol_rkpq()

    def inv_l(self) -> np.ndarray:
        return self.ctx.chol_inv_l()


def accuracy_to_tau(Accuracy: int) -> float:
    """Cholesky_Gammcor.f90:398-410."""
    try:
        return lib.accuracy_tau(Accuracy)
    except lib.GigpuError:
        raise GammcorError("Invalid accuracy setting")


def chol_gammcor_Rkpq(AOBasis, Accuracy: int, Omega: float | None = None, ctx: "lib.Context" = None,
                      decode_fixed: bool = False, params: TChol2Params | None = None,
                      time_eri: bool = False) -> TCholeskyVecsOTF:
    """Cholesky vectors of the Coulomb matrix in the AO basis computed on
    the fly (Cholesky_Gammcor.f90:335-441).  ``decode_fixed`` selects the
    Step-1 triangle order when Step 2 decodes same-shell pivots (the
    reference uses chol2_pq2p_ge_q, TwoStepCholesky_Step2.f90:527-528)."""
    if AOBasis.lmax > 5:
        raise GammcorError("Basis set includes angular momenta unsupported by the Auto2e subroutine")
    tau = accuracy_to_tau(Accuracy)
    kappa = 1.0 / Omega ** 2 if (Omega is not None and Omega > 0.0) else 0.0
    p = params or TChol2Params()
    p.CholeskyTauThresh, p.Kappa = tau, kappa
    if ctx is None:
        ctx = lib.Context(0)
    ctx.set_basis(AOBasis)
    flags = (lib.CHOL_DECODE_FIXED if decode_fixed else 0) | (lib.CHOL_TIME_ERI if time_eri else 0)
    try:
        ctx.chol_build(tau, kappa, p.MaxBlockDim, p.CholVecsBlock, flags)
    except lib.GigpuError as e:
        raise GammcorError(str(e))
    m = ctx.chol_meta()
    v = TChol2Vecs(NVecs=m["nvecs"], MaxSubsetDim=m["max_subset_dim"], NSubsets=(m["nsubsets"], 1),
                   NOrbPairs=m["norb_pairs"], NShellPairs=m["nshell_pairs"], ShellPairs=m["shell_pairs"],
                   ShellPairLoc=m["shell_pair_loc"], ShellPairDim=m["shell_pair_dim"], SubsetDim=m["subset_dim"],
                   SubsetBounds=m["subset_bounds"], Pivots=m["pivots"], PivotShellPairs=m["pivot_shell_pairs"],
                   PivotShellPairLoc=m["pivot_shell_pair_loc"], PivotShellPairDim=m["pivot_shell_pair_dim"],
                   PivotOrbPairs=m["pivot_orb_pairs"], NPivotShellPairs=m["npivot_shell_pairs"],
                   pivot_order=m["pivot_order"])
    return TCholeskyVecsOTF(Chol2Data=v, ctx=ctx, timings=ctx.chol_timings())


def auto2e_interface_AngFuncTransf(X_in: np.ndarray, AOBasis, ExternalOrdering: int) -> np.ndarray:
    """External-program AO order -> Auto2e order on the rows of X
    (Auto2eInterface.f90:217-275).  For spherical AOs the Dalton and PySCF
    orders equal the Auto2e order (s; px, py, pz; m = -l..l), so the map is
    the identity (Auto2eInterface.f90:68-101)."""
    if ExternalOrdering in (ORBITAL_ORDERING_AUTO2E,):
        return X_in
    if ExternalOrdering in (ORBITAL_ORDERING_DALTON, ORBITAL_ORDERING_PYSCF) and AOBasis.spher_ao:
        return X_in
    raise GammcorError("AO ordering %d is not supported by this build" % ExternalOrdering)


def chol_gammcor_Rkab(Rkab, CA, a0, a1, CB, b0, b1, MaxBufferDimMB, CholeskyVecsOTF: TCholeskyVecsOTF, AOBasis,
                      ExternalOrdering: int, all_columns: bool = False):
    """Rkab(NCholesky, NA*NB) of the MO-basis Cholesky vectors
    (Cholesky_Gammcor.f90:26-92).  ``Rkab`` may be None (allocated here) or
    an (NCholesky, NA*NB) array; the filled array is returned."""
    NA, NB = a1 - a0 + 1, b1 - b0 + 1
    nao = AOBasis.nao_spher if AOBasis.spher_ao else AOBasis.nao_cart
    CA = np.asarray(CA, dtype=np.float64)
    CB = np.asarray(CB, dtype=np.float64)
    if CA.shape[0] != nao or CB.shape[0] != nao:
        raise GammcorError("Invalid number of atomic orbitals in CA/CB")
    nchol = CholeskyVecsOTF.Chol2Data.NVecs
    if Rkab is not None and (Rkab.shape[0] != nchol or Rkab.shape[1] != NA * NB):
        raise GammcorError("Invalid dimension of Rkab")
    CA_ao = auto2e_interface_AngFuncTransf(CA, AOBasis, ExternalOrdering)
    CB_ao = auto2e_interface_AngFuncTransf(CB, AOBasis, ExternalOrdering)
    try:
        out = CholeskyVecsOTF.ctx.mo_transform(CA_ao, a0, a1, CB_ao, b0, b1, MaxBufferDimMB,
                                               lib.MO_ALL_COLUMNS if all_columns else 0)
    except lib.GigpuError as e:
        msg = str(e)
        if "Buffer size too small" in msg:
            raise GammcorError("Buffer size too small for Rpq->Rab")
        raise GammcorError(msg)
    if Rkab is not None:
        Rkab[...] = out
        return Rkab
    return out
