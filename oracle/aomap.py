"""Oracle of the external AO orderings (test infrastructure: only tests/
use it).  A plain-Python restatement of auto2e_interface_AOMap
(Auto2eInterface.f90:378-514) and auto2e_interface_ApplyOrcaPhases_Vector
(:517-562) with the reference's 1-based tables (:24-213): for external
position i of a shell of angular momentum l, table[l][i] is the 1-based
position of that function in the Auto2e order of the shell.

Pins: every map is a permutation; Dalton, PySCF and (per shell) Open Molcas
tables are the identity; the Molpro/Orca spherical tables list m values
whose offsets by l + 1 reproduce the tables; the Molcas map groups an
atom's functions by l, m, shell (tests/test_oracle_aomap.py).
"""
from __future__ import annotations

import numpy as np

AUTO2E, MOLPRO, ORCA, DALTON, OPEN_MOLCAS, PYSCF = range(6)


def _m_table(ms, l):
    # the reference writes the m values and adds l + 1 (:24-83)
    return [m + l + 1 for m in ms]


SPHER = {
    MOLPRO: {0: [1], 1: [1, 2, 3],
             2: _m_table([0, -2, 1, 2, -1], 2),
             3: _m_table([1, -1, 0, 3, -2, -3, 2], 3),
             4: _m_table([0, -2, 1, 4, -1, 2, -4, 3, -3], 4),
             5: _m_table([1, -1, 2, 3, -4, -3, 4, -5, 0, 5, -2], 5)},
    ORCA: {0: [1], 1: [3, 1, 2],
           2: _m_table([0, 1, -1, 2, -2], 2),
           3: _m_table([0, 1, -1, 2, -2, 3, -3], 3),
           4: _m_table([0, 1, -1, 2, -2, 3, -3, 4, -4], 4),
           5: _m_table([0, 1, -1, 2, -2, 3, -3, 4, -4, 5, -5], 5)},
}
for _o in (DALTON, PYSCF, OPEN_MOLCAS):
    SPHER[_o] = {0: [1], 1: [1, 2, 3]}
    for _l in range(2, 6):
        SPHER[_o][_l] = _m_table(list(range(-_l, _l + 1)), _l)

CART_MOLPRO = {0: [1], 1: [1, 2, 3], 2: [1, 4, 6, 2, 3, 5],
               3: [1, 7, 10, 2, 3, 4, 8, 6, 9, 5],
               4: [1, 11, 15, 2, 3, 7, 12, 10, 14, 4, 6, 13, 5, 8, 9],
               5: list(range(1, 22))}


class OrderingError(ValueError):
    pass


def ao_map(shell_l, shell_atom, spher, ordering):
    """(Map_extao2ao 1-based, phase in Auto2e order) for shells given by their
    angular momenta and (0-based) atoms, shells of an atom consecutive."""
    ns = len(shell_l)
    nang = [(2 * l + 1) if spher else (l + 1) * (l + 2) // 2 for l in shell_l]
    loc = [1]
    for k in range(ns - 1):
        loc.append(loc[-1] + nang[k])
    nao = sum(nang)
    Map = [-1] * (nao + 1)             # 1-based, Map[0] unused
    if spher:
        if ordering not in SPHER:
            raise OrderingError("Invalid external orbital ordering")
        table = SPHER[ordering]
        if ordering == OPEN_MOLCAS:
            s0 = 0
            while s0 < ns:
                s1 = s0
                while s1 + 1 < ns and shell_atom[s1 + 1] == shell_atom[s0]:
                    s1 += 1
                i = loc[s0]
                for L in range(0, max(shell_l[s0:s1 + 1]) + 1):
                    for M in range(1, 2 * L + 2):
                        for s in range(s0, s1 + 1):
                            if shell_l[s] == L:
                                Map[i] = loc[s] + table[L][M - 1] - 1
                                i += 1
                s0 = s1 + 1
        else:
            for kk in range(ns):
                L = shell_l[kk]
                for p in range(loc[kk], loc[kk] + nang[kk]):
                    Map[p] = loc[kk] + table[L][p - loc[kk]] - 1
    else:
        if ordering != MOLPRO:
            raise OrderingError("Invalid external orbital ordering")
        for kk in range(ns):
            L = shell_l[kk]
            for p in range(loc[kk], loc[kk] + nang[kk]):
                Map[p] = loc[kk] + CART_MOLPRO[L][p - loc[kk]] - 1
    phase = np.ones(nao)
    if spher and ordering == ORCA:
        for kk in range(ns):
            L = shell_l[kk]
            flips = {3: (3,), 4: (3, 4), 5: (3, 4)}.get(L, ())
            for d in flips:
                for p in (loc[kk] + L + d, loc[kk] + L - d):     # 1-based, as the reference
                    phase[p - 1] = -phase[p - 1]
    return np.array(Map[1:], dtype=np.int64), phase


def basis_shells(basis):
    """(shell_l, shell_atom) of a gigpu AOBasis."""
    ls = [basis.shell_l(s) for s in range(basis.nshells)]
    atoms = [int(basis.shell_centers[s]) - 1 for s in range(basis.nshells)]
    return ls, atoms
