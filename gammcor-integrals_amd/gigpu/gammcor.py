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


def auto2e_interface_AOMap(AOBasis, ExternalOrdering: int, ctx):
    """(Map_extao2ao, Map_ao2extao), 1-based (auto2e_interface_AOMap,
    Auto2eInterface.f90:378-514), computed by the library for the basis set
    on ``ctx`` (gi_ao_map).  Raises GammcorError("Invalid external orbital
    ordering") where the reference stops."""
    try:
        m, _ = ctx.ao_map(ExternalOrdering)
    except lib.GigpuError as e:
        if "gigpu error 70:" in str(e):
            raise GammcorError("Invalid external orbital ordering")
        raise
    inv = np.empty_like(m)
    inv[m - 1] = np.arange(1, len(m) + 1)
    return m, inv


def auto2e_interface_AngFuncTransf(X_in: np.ndarray, AOBasis, ExternalOrdering: int, FromExternalAO: bool = True,
                                   TwoIndexTransf: bool = False, ctx=None) -> np.ndarray:
    """X_out from X_in between the external program's AO order and the Auto2e
    order (Auto2eInterface.f90:217-275): FromExternalAO -> Map_extao2ao, else
    Map_ao2extao; X_out(Map(p), :) = X_in(p, :), and with TwoIndexTransf also
    the columns, X_out(Map(p), Map(q)) = X_in(p, q) (:278-310).  No signs: the
    callers apply the Orca phases (auto2e_interface_ApplyOrcaPhases_Matrix)."""
    ctx = _context_for(AOBasis, ctx)
    ext2ao, ao2ext = auto2e_interface_AOMap(AOBasis, ExternalOrdering, ctx)
    Map = (ext2ao if FromExternalAO else ao2ext) - 1
    X_in = np.asarray(X_in, dtype=np.float64)
    X_out = np.empty_like(X_in)
    X_out[Map, ...] = X_in
    if TwoIndexTransf:
        W = X_out.copy()
        X_out[:, Map] = W
    return X_out


def auto2e_interface_ApplyOrcaPhases_Matrix(X: np.ndarray, AOBasis, TwoIndexTransf: bool, ctx=None) -> np.ndarray:
    """Orca's signs of the spherical functions with m = +-3 (f, g, h) and
    m = +-4 (g, h) (Auto2eInterface.f90:517-585), applied in place to the
    rows (and with TwoIndexTransf the columns) of X at their Auto2e
    positions.  The reference's row loop runs over NAO columns (:574-577);
    here every column of X is covered, which is the same for the square
    matrices it is called with."""
    ctx = _context_for(AOBasis, ctx)
    _, phase = ctx.ao_map(ORBITAL_ORDERING_ORCA)
    X *= phase[:, None]
    if TwoIndexTransf:
        X *= phase[None, :]
    return X


def auto2e_interface_C(C_extao: np.ndarray, AOBasis, ExternalOrdering: int, ctx=None) -> np.ndarray:
    """C(external AO order) -> C(Auto2e order), Orca phases included
    (auto2e_interface_C, Auto2eInterface.f90:313-337)."""
    C_ao = auto2e_interface_AngFuncTransf(C_extao, AOBasis, ExternalOrdering, True, False, ctx)
    if ExternalOrdering == ORBITAL_ORDERING_ORCA:
        auto2e_interface_ApplyOrcaPhases_Matrix(C_ao, AOBasis, False, ctx)
    return C_ao


def auto2e_interface_V(V_extao: np.ndarray, AOBasis, ExternalOrdering: int, ctx=None) -> np.ndarray:
    """V(external AO order, both indices) -> V(Auto2e order)
    (auto2e_interface_V, Auto2eInterface.f90:340-361)."""
    V_ao = auto2e_interface_AngFuncTransf(V_extao, AOBasis, ExternalOrdering, True, True, ctx)
    if ExternalOrdering == ORBITAL_ORDERING_ORCA:
        auto2e_interface_ApplyOrcaPhases_Matrix(V_ao, AOBasis, True, ctx)
    return V_ao


def _context_for(AOBasis, ctx):
    """A library context holding AOBasis (the AO maps are computed there)."""
    if ctx is not None and getattr(ctx, "basis", None) is AOBasis:
        return ctx
    global _map_ctx
    if _map_ctx is None:
        _map_ctx = lib.Context(0)
    if getattr(_map_ctx, "basis", None) is not AOBasis:
        _map_ctx.set_basis(AOBasis)
    return _map_ctx


_map_ctx = None


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
    # external AO order -> Auto2e order, Orca phases (Cholesky_Gammcor.f90:73-83)
    ctx = CholeskyVecsOTF.ctx
    CA_ao = auto2e_interface_C(CA, AOBasis, ExternalOrdering, ctx)
    CB_ao = auto2e_interface_C(CB, AOBasis, ExternalOrdering, ctx)
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
