"""Host-side mirror of the reference's ``OneElectronInts_Gammcor`` module
(src/gammcor-interface/OneElectronInts_Gammcor.f90), running on libgigpu.

* ``ints1e_gammcor_H0(H0_ao, AOBasis, System)`` (:13-49)
* ``ints1e_gammcor_H0_mo(H0_mo, C, AOBasis, System, ExternalOrdering)`` (:76-126)
* ``ints1e_gammcor_H0_extao(H0_extao, AOBasis, System, ExternalOrdering)`` (:129-162)

``System`` carries what the reference reads from TSystem for the nuclear
attraction (ints1e_Coulomb, OneElectronInts.f90:518-712): the nuclear charges
of the real atoms (ZNumbers, or ZNumbersECP when ECPCharges) and the external
point charges.  Errors raise ``GammcorError`` with the reference's messages.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import lib
from .gammcor import (GammcorError, ORBITAL_ORDERING_ORCA, auto2e_interface_AngFuncTransf,
                      auto2e_interface_ApplyOrcaPhases_Matrix, auto2e_interface_C)


@dataclass
class TSystem:
    ZNumbers: np.ndarray                          # (natoms,) nuclear charges, 0 for dummy atoms
    ECPCharges: bool = False
    ZNumbersECP: np.ndarray | None = None         # charges left by effective core potentials
    PointCharges: np.ndarray = field(default_factory=lambda: np.zeros(0))
    PointChargeCoords: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))   # bohr

    def atom_charges(self) -> np.ndarray:
        z = self.ZNumbersECP if self.ECPCharges else self.ZNumbers
        return np.asarray(z, dtype=np.float64)

    def point_charges(self):
        q = np.asarray(self.PointCharges, dtype=np.float64).reshape(-1)
        if len(q) == 0:
            return None
        xyz = np.asarray(self.PointChargeCoords, dtype=np.float64).reshape(-1, 3)
        return np.column_stack([xyz, q])


_ELEMENT_Z = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}


def system_of(AOBasis) -> TSystem:
    """TSystem with the nuclear charges of the basis' atom symbols."""
    return TSystem(ZNumbers=np.array([_ELEMENT_Z.get(s, 0) for s in AOBasis.atom_symbols], dtype=np.float64))


def _context(AOBasis, ctx):
    if ctx is None:
        ctx = lib.Context(0)
    ctx.set_basis(AOBasis)
    return ctx


def ints1e_gammcor_H0(H0_ao, AOBasis, System: TSystem, ctx: "lib.Context" = None):
    """Bare-nucleus Hamiltonian T + V_ne in the AO basis (Auto2e order)."""
    nao = AOBasis.nao_spher if AOBasis.spher_ao else AOBasis.nao_cart
    if H0_ao is not None and H0_ao.shape[0] != nao:
        raise GammcorError("Invalid dimension of matrix C_ao")
    ctx = _context(AOBasis, ctx)
    H = ctx.h0(System.atom_charges(), System.point_charges())
    if H0_ao is not None:
        H0_ao[...] = H
        return H0_ao
    return H


def ints1e_gammcor_H0_mo(H0_mo, C, AOBasis, System: TSystem, ExternalOrdering: int, ctx: "lib.Context" = None):
    """C_ao^T H0 C_ao, C given in the external program's AO order."""
    C = np.asarray(C, dtype=np.float64)
    nao = AOBasis.nao_spher if AOBasis.spher_ao else AOBasis.nao_cart
    if C.shape[0] != nao:
        raise GammcorError("Invalid dimension of matrix C")
    ctx = _context(AOBasis, ctx)
    H0_ao = ints1e_gammcor_H0(None, AOBasis, System, ctx)
    # MO coefficients in the Auto2e order, Orca phases included
    # (OneElectronInts_Gammcor.f90:107-113)
    C_ao = auto2e_interface_C(C, AOBasis, ExternalOrdering, ctx)
    out = C_ao.T @ (H0_ao @ C_ao)
    if H0_mo is not None:
        H0_mo[...] = out
        return H0_mo
    return out


def ints1e_gammcor_H0_extao(H0_extao, AOBasis, System: TSystem, ExternalOrdering: int,
                            ctx: "lib.Context" = None):
    """H0 with both indices in the external program's AO order."""
    nao = AOBasis.nao_spher if AOBasis.spher_ao else AOBasis.nao_cart
    if H0_extao is not None and H0_extao.shape[0] != nao:
        raise GammcorError("Invalid dimension of matrix H0_extao")
    ctx = _context(AOBasis, ctx)
    H0_ao = ints1e_gammcor_H0(None, AOBasis, System, ctx)
    # Auto2e -> external order on both indices, then the Orca signs applied
    # at their Auto2e positions to the externally ordered matrix, as the
    # reference does (OneElectronInts_Gammcor.f90:152-159)
    out = auto2e_interface_AngFuncTransf(H0_ao, AOBasis, ExternalOrdering, False, True, ctx)
    if ExternalOrdering == ORBITAL_ORDERING_ORCA:
        auto2e_interface_ApplyOrcaPhases_Matrix(out, AOBasis, True, ctx)
    if H0_extao is not None:
        H0_extao[...] = out
        return H0_extao
    return out
