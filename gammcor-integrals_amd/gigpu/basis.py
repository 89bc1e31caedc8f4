"""Gaussian AO basis in the layout of the reference's ``TAOBasis``.

Reference: ``src/basis/basis_definitions.f90:94-192`` (the type),
``src/basis/basis_sets.f90:826-855`` (``basis_CountOrbitals``),
``src/basis/basis_sets.f90:894-944`` (``basis_NormalizeCntrCoeffs``) and
``src/basis/sphergto.f90:184-268`` (``SpherGTO_NormFactors_Shell`` /
``ShellOverlap``).  Arrays keep the reference's 1-based index values
(shell centres, shell parameter sets) so they can be handed to the C ABI
unchanged, exactly as the Fortran shim would pass ``AOBasis`` components.

A contracted AO is ``N * S(x,y,z) * sum_i c_i exp(-alpha_i r^2)`` where
``S`` is ``x^lx y^ly z^lz`` (Cartesian) or a real solid harmonic normalized
on the unit sphere (spherical, l >= 2; l = 0, 1 stay s and x, y, z), ``c_i``
carry the primitive normalization and ``N`` makes the AO unit-normalized.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

PI = 3.141592653589793          # math_constants.f90:6
BOHR_ANGSTROM = 0.52917721092   # math_constants.f90:160 (tobohr divides by it)


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def cart_components(l: int):
    """(lx, ly, lz) in the order lx = l..0, ly = l-lx..0
    (Auto2eGlobal.py:35-43, basis_sets.f90:797-807)."""
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def dblfact(n: int) -> float:
    """(n)!! with (-1)!! = 1."""
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def spher_u(l: int) -> np.ndarray:
    """Cartesian -> real solid harmonic coefficients, shape (2l+1, ncart(l)),
    rows m = -l..l.  Helgaker eqs. 6.4.47-50 rescaled to unit-sphere
    normalization, as in AutoSpherTransf.py:62-158 (Nslm, Clmtuv,
    RSHU_l_eq_j)."""
    comps = cart_components(l)
    index = {c: i for i, c in enumerate(comps)}
    U = np.zeros((2 * l + 1, len(comps)))
    for m in range(-l, l + 1):
        am = abs(m)
        vm = 0 if m >= 0 else 1
        npref = (math.sqrt((2 * l + 1) / (4 * PI)) * math.sqrt(math.factorial(l + am))
                 * math.sqrt(math.factorial(l - am)) * (1.0 if m == 0 else math.sqrt(2.0))
                 / (2 ** am * math.factorial(l)))
        for t in range((l - am) // 2 + 1):
            for u in range(t + 1):
                for k in range((am - vm) // 2 + 1):
                    v = 2 * k + vm
                    lx = 2 * t + am - 2 * u - v
                    ly = 2 * u + v
                    lz = l - 2 * t - am
                    a = math.comb(l, t) * (math.comb(l - t, am + t) if am + t <= l - t else 0)
                    b = math.comb(t, u) * math.comb(am, v)
                    c = (-1) ** (t + (v - vm) // 2) * 4 ** t
                    U[m + l, index[(lx, ly, lz)]] += npref * (a * b / c)
    return U


def shell_overlap_cart(l: int, cntr: np.ndarray, exps: np.ndarray) -> np.ndarray:
    """Overlap of the (unnormalized) Cartesian components of one contracted
    shell with itself (sphergto.f90:214-268)."""
    comps = cart_components(l)
    n = len(comps)
    S = np.zeros((n, n))
    for k in range(len(exps)):
        for q in range(len(exps)):
            ap = exps[k] + exps[q]
            one_d = {}
            for t in range(0, l + 1):
                one_d[2 * t] = dblfact(2 * t - 1) / (2.0 * ap) ** t * math.sqrt(PI / ap)
            for bi, cb in enumerate(comps):
                for ai, ca in enumerate(comps):
                    sx, sy, sz = ca[0] + cb[0], ca[1] + cb[1], ca[2] + cb[2]
                    if sx % 2 or sy % 2 or sz % 2:
                        continue
                    S[ai, bi] += cntr[k] * cntr[q] * one_d[sx] * one_d[sy] * one_d[sz]
    return S


def normalize_shell(l: int, exps: np.ndarray, coeffs: np.ndarray):
    """Primitive-normalize the contraction coefficients and return
    (cntr, norm_cart, norm_spher) for one shell parameter set
    (basis_sets.f90:894-944, sphergto.f90:184-211)."""
    exps = np.asarray(exps, dtype=np.float64)
    cntr = np.array(coeffs, dtype=np.float64)
    for i in range(len(exps)):
        cntr[i] = cntr[i] * (2.0 / PI) ** (3.0 / 4.0) * 2.0 ** l * exps[i] ** ((2.0 * l + 3.0) / 4.0)
    z = 0.0
    for i in range(len(exps)):
        for j in range(len(exps)):
            z += cntr[i] * cntr[j] / (exps[i] + exps[j]) ** (l + 3.0 / 2.0)
    y = math.sqrt(PI ** (3.0 / 2.0) / 2.0 ** l)
    z = math.sqrt(z)
    norm_cart = np.array([1.0 / (math.sqrt(dblfact(2 * lx - 1) * dblfact(2 * ly - 1)
                                           * dblfact(2 * lz - 1)) * y * z)
                          for (lx, ly, lz) in cart_components(l)])
    if l >= 2:
        U = spher_u(l)
    else:
        U = np.eye(ncart(l))
    Scc = shell_overlap_cart(l, cntr, exps)
    Tss = U @ Scc @ U.T
    norm_spher = 1.0 / np.sqrt(np.diag(Tss))
    return cntr, norm_cart, norm_spher


@dataclass
class AOBasis:
    """SoA basis arrays mirroring TAOBasis (basis_definitions.f90:94-192).

    Index-valued arrays hold 1-based values, as in the reference."""
    atom_coords: np.ndarray          # (natoms, 3) bohr; memory = Fortran (3, natoms)
    atom_symbols: list
    shell_centers: np.ndarray        # (nshells,) int64, 1-based atoms
    shell_params_idx: np.ndarray     # (nshells,) int64, 1-based params
    shell_momentum: np.ndarray       # (nparams,) int64
    nprimitives: np.ndarray          # (nparams,) int64
    cntr_coeffs: np.ndarray          # (nparams, max_nprim): Fortran (max_nprim, nparams)
    exponents: np.ndarray            # (nparams, max_nprim)
    norm_cart: np.ndarray            # (nparams, max_ncart)
    norm_spher: np.ndarray           # (nparams, max_nspher)
    spher_ao: bool = True
    shell_loc_spher: np.ndarray = field(default=None)
    shell_loc_cart: np.ndarray = field(default=None)

    def __post_init__(self):
        nsph = np.array([2 * self.shell_momentum[p - 1] + 1 for p in self.shell_params_idx], np.int64)
        ncar = np.array([ncart(int(self.shell_momentum[p - 1])) for p in self.shell_params_idx], np.int64)
        self.shell_loc_spher = (np.concatenate([[0], np.cumsum(nsph)[:-1]]) + 1).astype(np.int64)
        self.shell_loc_cart = (np.concatenate([[0], np.cumsum(ncar)[:-1]]) + 1).astype(np.int64)
        self.nao_spher = int(nsph.sum())
        self.nao_cart = int(ncar.sum())

    # ---- convenience -----------------------------------------------------
    @property
    def natoms(self) -> int:
        return self.atom_coords.shape[0]

    @property
    def nshells(self) -> int:
        return len(self.shell_centers)

    @property
    def nparams(self) -> int:
        return len(self.shell_momentum)

    @property
    def lmax(self) -> int:
        return int(self.shell_momentum.max())

    @property
    def nao(self) -> int:
        return self.nao_spher if self.spher_ao else self.nao_cart

    @property
    def norm_factors(self) -> np.ndarray:
        """(nparams, max_nang) of the active AO kind."""
        return self.norm_spher if self.spher_ao else self.norm_cart

    def shell_nang(self, s: int) -> int:
        l = int(self.shell_momentum[self.shell_params_idx[s] - 1])
        return 2 * l + 1 if self.spher_ao else ncart(l)

    def shell_l(self, s: int) -> int:
        return int(self.shell_momentum[self.shell_params_idx[s] - 1])

    def shell_loc(self) -> np.ndarray:
        return self.shell_loc_spher if self.spher_ao else self.shell_loc_cart

    def c_arrays(self) -> dict:
        """Contiguous arrays in the exact memory layout of the Fortran
        allocatables (column-major), ready for the C ABI."""
        return dict(
            coords=np.ascontiguousarray(self.atom_coords, dtype=np.float64),
            shell_centers=np.ascontiguousarray(self.shell_centers, dtype=np.int64),
            shell_params=np.ascontiguousarray(self.shell_params_idx, dtype=np.int64),
            momentum=np.ascontiguousarray(self.shell_momentum, dtype=np.int64),
            nprim=np.ascontiguousarray(self.nprimitives, dtype=np.int64),
            cntr=np.ascontiguousarray(self.cntr_coeffs, dtype=np.float64),
            exps=np.ascontiguousarray(self.exponents, dtype=np.float64),
            norm=np.ascontiguousarray(self.norm_factors, dtype=np.float64),
        )


def build_basis(symbols, coords_bohr, element_shells, spher_ao=True) -> AOBasis:
    """Assemble an AOBasis with SHELL_ORDER_FIXED: for every atom, its
    element's shells in the given order (basis_sets.f90:316-385).

    element_shells: {symbol: [(l, [exponents], [coefficients]), ...]}"""
    elements = []
    for s in symbols:
        if s not in elements:
            elements.append(s)
    params = []          # (l, exps, cntr, ncart_norm, nspher_norm)
    param_index = {}
    for el in elements:
        for k, (l, exps, coefs) in enumerate(element_shells[el]):
            cntr, ncn, nsn = normalize_shell(l, np.asarray(exps, float), np.asarray(coefs, float))
            param_index[(el, k)] = len(params) + 1
            params.append((l, np.asarray(exps, float), cntr, ncn, nsn))
    max_nprim = max(len(p[1]) for p in params)
    lmax = max(p[0] for p in params)
    nparams = len(params)
    mom = np.array([p[0] for p in params], np.int64)
    nprim = np.array([len(p[1]) for p in params], np.int64)
    cntr = np.zeros((nparams, max_nprim))
    exps = np.zeros((nparams, max_nprim))
    ncn = np.zeros((nparams, ncart(lmax)))
    nsn = np.zeros((nparams, 2 * lmax + 1))
    for i, (l, e, c, a, b) in enumerate(params):
        exps[i, :len(e)] = e
        cntr[i, :len(c)] = c
        ncn[i, :len(a)] = a
        nsn[i, :len(b)] = b
    centers, pidx = [], []
    for ia, s in enumerate(symbols):
        for k in range(len(element_shells[s])):
            centers.append(ia + 1)
            pidx.append(param_index[(s, k)])
    return AOBasis(atom_coords=np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3),
                   atom_symbols=list(symbols),
                   shell_centers=np.array(centers, np.int64),
                   shell_params_idx=np.array(pidx, np.int64),
                   shell_momentum=mom, nprimitives=nprim, cntr_coeffs=cntr,
                   exponents=exps, norm_cart=ncn, norm_spher=nsn, spher_ao=spher_ao)
