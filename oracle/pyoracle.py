"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product library.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

i64 = C.c_int64
dptr = C.POINTER(C.c_double)
iptr = C.POINTER(C.c_int64)


class OrBasis(C.Structure):
    _fields_ = [("natoms", i64), ("coords", dptr), ("nshells", i64), ("shell_centers", iptr),
                ("shell_params", iptr), ("nparams", i64), ("momentum", iptr), ("nprim", iptr),
                ("max_nprim", i64), ("cntr", dptr), ("exps", dptr), ("max_nang", i64),
                ("norm", dptr), ("spher", i64)]


class OrCholResult(C.Structure):
    _fields_ = [("status", i64), ("message", C.c_char * 256), ("nvecs", i64), ("norb_pairs", i64),
                ("nshell_pairs", i64), ("nsubsets", i64), ("max_subset_dim", i64),
                ("npivot_shell_pairs", i64), ("nmacro", i64),
                ("shell_pairs", iptr), ("shell_pair_loc", iptr), ("shell_pair_dim", iptr),
                ("subset_dim", iptr), ("subset_bounds", iptr), ("pivots", iptr),
                ("pivot_order", iptr), ("pivot_shell_pairs", iptr), ("pivot_shell_pair_loc", iptr),
                ("pivot_shell_pair_dim", iptr), ("pivot_orb_pairs", iptr), ("vdiag", dptr),
                ("inv_l", dptr), ("rkpq", dptr), ("max_residual", C.c_double),
                ("t_step1", C.c_double), ("t_step2", C.c_double), ("t_vectors", C.c_double),
                ("t_2e", C.c_double), ("t_rtr", C.c_double), ("t_rk", C.c_double),
                ("t_compress", C.c_double), ("t_vpqrs", C.c_double), ("t_chol", C.c_double),
                ("t_inv", C.c_double), ("nquartets", i64)]


DECODE_FIXED = 1
SKIP_VECTORS = 2
RAW_TIES = 4

_lib = None


def build(force: bool = False) -> None:
    src = [os.path.join(HERE, f) for f in ("boys.c", "eri.c", "sort.c", "linalg.c", "chol.c", "mo.c", "coulomb.c",
                                           "onee.c", "oracle.h")]
    if (not force and os.path.exists(LIB_PATH)
            and os.path.getmtime(LIB_PATH) >= max(os.path.getmtime(s) for s in src)):
        return
    subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_boys_init.argtypes = [i64]
        L.or_boys_fm_reference.argtypes = [i64, C.c_double]
        L.or_boys_fm_reference.restype = C.c_double
        L.or_boys_fm.argtypes = [i64, C.c_double, dptr]
        L.or_boys_table.argtypes = [iptr, iptr]
        L.or_boys_table.restype = dptr
        L.or_spher_u.argtypes = [i64, dptr]
        L.or_nang.argtypes = [C.POINTER(OrBasis), i64]
        L.or_nang.restype = i64
        L.or_eri_quartet.argtypes = [C.POINTER(OrBasis), i64, i64, i64, i64, C.c_double, dptr]
        L.or_eri_quartet_single_boys.argtypes = [C.POINTER(OrBasis), i64, i64, i64, i64, C.c_double, dptr]
        L.or_one_electron.argtypes = [C.POINTER(OrBasis), dptr, i64, dptr, dptr, C.c_void_p, C.c_void_p,
                                      C.c_void_p, i64]
        L.or_coulomb_energy.argtypes = [C.POINTER(OrBasis), dptr, i64, C.c_double]
        L.or_coulomb_energy.restype = C.c_double
        L.or_eri_deriv_a.argtypes = [C.POINTER(OrBasis), i64, i64, i64, i64, C.c_double, dptr]
        L.or_eri_deriv_a.restype = C.c_int
        L.or_coulomb_grad.argtypes = [C.POINTER(OrBasis), dptr, i64, C.c_double, i64, iptr, dptr]
        L.or_coulomb_grad.restype = C.c_int
        L.or_eri_pair_matrix.argtypes = [C.POINTER(OrBasis), i64, iptr, C.c_double, dptr, i64]
        L.or_dsort0.argtypes = [dptr, iptr, i64, i64]
        L.or_chol_build.argtypes = [C.POINTER(OrBasis), C.c_double, C.c_double, i64, i64, i64,
                                    C.POINTER(OrCholResult)]
        L.or_chol_free.argtypes = [C.POINTER(OrCholResult)]
        L.or_pq2p_ge_q.argtypes = [i64, i64, iptr, iptr]
        L.or_mo_transform.argtypes = [C.POINTER(OrBasis), C.POINTER(OrCholResult), i64, dptr, i64, i64, i64,
                                      dptr, i64, i64, i64, i64, dptr]
        L.or_mo_transform.restype = i64
        L.or_dpotrf_lower.argtypes = [i64, dptr, i64]
        L.or_dpotrf_lower.restype = i64
        L.or_dtrtri_lower.argtypes = [i64, dptr, i64]
        L.or_dtrtri_lower.restype = i64
        L.or_dgemm.argtypes = [C.c_char, C.c_char, i64, i64, i64, C.c_double, dptr, i64, dptr, i64,
                               C.c_double, dptr, i64]
        L.or_boys_init(20)
        _lib = L
    return _lib


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    """Oracle bound to one AOBasis (keeps the numpy arrays alive)."""

    def __init__(self, basis):
        self.basis = basis
        self.arr = basis.c_arrays()
        a = self.arr
        self.ob = OrBasis(basis.natoms, _p(a["coords"]), basis.nshells, _p(a["shell_centers"], i64),
                          _p(a["shell_params"], i64), basis.nparams, _p(a["momentum"], i64),
                          _p(a["nprim"], i64), a["cntr"].shape[1], _p(a["cntr"]), _p(a["exps"]),
                          a["norm"].shape[1], _p(a["norm"]), 1 if basis.spher_ao else 0)
        self.L = lib()

    def eri(self, sa, sb, sc, sd, kappa=0.0, single_boys=False):
        """(ab|cd) block V[a, b, c, d] for 0-based shells (numpy C-order
        indexing of the Fortran V(a,b,c,d))."""
        n = [self.L.or_nang(C.byref(self.ob), s) for s in (sa, sb, sc, sd)]
        V = np.zeros(n[0] * n[1] * n[2] * n[3])
        f = self.L.or_eri_quartet_single_boys if single_boys else self.L.or_eri_quartet
        f(C.byref(self.ob), sa, sb, sc, sd, kappa, _p(V))
        return V.reshape(n[3], n[2], n[1], n[0]).transpose(3, 2, 1, 0)

    def pair_matrix(self, shell_pairs, kappa=0.0):
        """Dense (pq|rs) over the orbital pairs of shell_pairs (1-based
        (A, B), A >= B), each pair's orbital pairs in the reference's storage
        order (b outer, a inner, a >= b for A == B)."""
        sp = np.ascontiguousarray(np.asarray(shell_pairs, dtype=np.int64).reshape(-1, 2) - 1)
        dims = []
        for a, b in sp:
            na, nb = self.basis.shell_nang(int(a)), self.basis.shell_nang(int(b))
            dims.append(na * (na + 1) // 2 if a == b else na * nb)
        n = int(sum(dims))
        V = np.zeros((n, n), order="F")
        self.L.or_eri_pair_matrix(C.byref(self.ob), len(sp), _p(sp, i64), float(kappa), _p(V), n)
        return V

    def one_electron(self, charges, point_charges=None):
        """(S, T, V) in the AO basis; charges per atom, point_charges an
        iterable of (x, y, z, q) in bohr."""
        n = self.basis.nao
        S, T, V = (np.zeros((n, n), order="F") for _ in range(3))
        q = np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
        pc = np.zeros((0, 4)) if point_charges is None else np.asarray(point_charges, dtype=np.float64).reshape(-1, 4)
        xyz = np.ascontiguousarray(pc[:, :3]) if len(pc) else np.zeros(3)
        qq = np.ascontiguousarray(pc[:, 3]) if len(pc) else np.zeros(1)
        self.L.or_one_electron(C.byref(self.ob), _p(q), len(pc), _p(xyz), _p(qq), S.ctypes.data, T.ctypes.data,
                               V.ctypes.data, n)
        return S, T, V

    def coulomb_energy(self, density, kappa=0.0):
        """E_J = 1/2 sum D_pq (pq|rs) D_rs of a symmetric AO density."""
        D = np.asfortranarray(np.asarray(density, dtype=np.float64))
        return float(self.L.or_coulomb_energy(C.byref(self.ob), _p(D), D.shape[0], kappa))

    def eri_deriv(self, sa, sb, sc, sd, kappa=0.0):
        """d(ab|cd)/dA for the centre A of shell sa (0-based shells; only
        shell sa's functions move): array [3, a, b, c, d]."""
        n = [self.L.or_nang(C.byref(self.ob), s) for s in (sa, sb, sc, sd)]
        dV = np.zeros(3 * n[0] * n[1] * n[2] * n[3])
        if self.L.or_eri_deriv_a(C.byref(self.ob), sa, sb, sc, sd, kappa, _p(dV)) != 0:
            raise ValueError("derivative class beyond the oracle's tables")
        return dV.reshape(3, n[3], n[2], n[1], n[0]).transpose(0, 4, 3, 2, 1)

    def coulomb_grad(self, density, kappa=0.0, atoms=None):
        """dE_J/dR (natoms, 3) from derivative ERIs; with `atoms` (0-based
        indices) only those rows are computed, the others are zero."""
        D = np.asfortranarray(np.asarray(density, dtype=np.float64))
        G = np.zeros(3 * self.basis.natoms)
        sel = np.ascontiguousarray(np.asarray([] if atoms is None else atoms, dtype=np.int64))
        nsel = -1 if atoms is None else len(sel)
        if self.L.or_coulomb_grad(C.byref(self.ob), _p(D), D.shape[0], kappa, nsel,
                                  _p(sel if len(sel) else np.zeros(1, np.int64), i64), _p(G)) != 0:
            raise ValueError("derivative class beyond the oracle's tables")
        return G.reshape(self.basis.natoms, 3)

    def chol(self, tau=1e-6, kappa=0.0, max_block_dim=4000, vec_block=500, flags=0):
        res = OrCholResult()
        self.L.or_chol_build(C.byref(self.ob), tau, kappa, max_block_dim, vec_block, flags, C.byref(res))
        return CholResult(self, res)


def _arr(ptr, n, dtype):
    if not ptr or n <= 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


class CholResult:
    def __init__(self, oracle, res):
        self.oracle = oracle
        self.res = res
        r = res
        self.status = r.status
        self.message = r.message.decode()
        self.nvecs = r.nvecs
        self.norb_pairs = r.norb_pairs
        self.nshell_pairs = r.nshell_pairs
        self.nsubsets = r.nsubsets
        self.max_subset_dim = r.max_subset_dim
        self.nmacro = r.nmacro
        self.nquartets = r.nquartets
        self.max_residual = r.max_residual
        nsp = r.nshell_pairs
        self.shell_pairs = _arr(r.shell_pairs, 2 * nsp, np.int64).reshape(nsp, 2)
        self.shell_pair_loc = _arr(r.shell_pair_loc, 3 * nsp, np.int64).reshape(nsp, 3)
        self.shell_pair_dim = _arr(r.shell_pair_dim, nsp, np.int64)
        self.subset_dim = _arr(r.subset_dim, r.nsubsets, np.int64)
        self.subset_bounds = _arr(r.subset_bounds, 2 * r.nsubsets, np.int64).reshape(r.nsubsets, 2)
        self.pivots = _arr(r.pivots, r.norb_pairs, np.int64)
        self.pivot_order = _arr(r.pivot_order, r.nvecs, np.int64)
        self.pivot_orb_pairs = _arr(r.pivot_orb_pairs, r.nvecs, np.int64)
        npsp = r.npivot_shell_pairs
        self.pivot_shell_pairs = _arr(r.pivot_shell_pairs, npsp, np.int64)
        self.pivot_shell_pair_loc = _arr(r.pivot_shell_pair_loc, npsp, np.int64)
        self.pivot_shell_pair_dim = _arr(r.pivot_shell_pair_dim, npsp, np.int64)
        nao = oracle.basis.nao
        self.vdiag = _arr(r.vdiag, nao * (nao + 1) // 2, np.float64)
        self.inv_l = _arr(r.inv_l, r.nvecs * r.nvecs, np.float64).reshape(r.nvecs, r.nvecs).T \
            if r.status == 0 or r.inv_l else None
        if r.rkpq:
            n = r.nvecs * r.max_subset_dim * r.nsubsets
            self.rkpq = _arr(r.rkpq, n, np.float64).reshape(r.nsubsets, r.max_subset_dim, r.nvecs)
        else:
            self.rkpq = None
        self.timings = dict(step1=r.t_step1, step2=r.t_step2, vectors=r.t_vectors, eri_step1=r.t_2e,
                            rtr=r.t_rtr, rk=r.t_rk, compress=r.t_compress, vpqrs=r.t_vpqrs,
                            cholesky=r.t_chol, inverse=r.t_inv)

    def set_rkpq(self, rkpq):
        """Replace the vectors the oracle's MO transform reads (same shape as
        self.rkpq), so that the transform can be checked on vectors produced
        elsewhere."""
        r = self.res
        n = r.nvecs * r.max_subset_dim * r.nsubsets
        rk = np.ascontiguousarray(rkpq, dtype=np.float64).reshape(-1)
        if not r.rkpq or rk.size != n:
            raise ValueError("rkpq shape does not match this decomposition")
        np.ctypeslib.as_array(r.rkpq, shape=(n,))[:] = rk
        self.rkpq = rk.reshape(r.nsubsets, r.max_subset_dim, r.nvecs).copy()

    def full_vectors(self):
        """Cholesky vectors as a dense (nvecs, norb_pairs) array in FULL
        storage order (see full_vectors())."""
        return full_vectors(self.shell_pair_loc, self.shell_pair_dim, self.rkpq, self.norb_pairs)

    def mo_transform(self, CA, a0, a1, CB, b0, b1, max_buffer_mb=100000):
        """Rkab (nvecs, NA*NB) in Fortran layout, returned as (NA*NB, nvecs)
        C-order view transposed back to (nvecs, NA*NB)."""
        nao = self.oracle.basis.nao
        CA = np.asfortranarray(CA, dtype=np.float64)
        CB = np.asfortranarray(CB, dtype=np.float64)
        NA, NB = a1 - a0 + 1, b1 - b0 + 1
        out = np.zeros(self.nvecs * NA * NB)
        st = self.oracle.L.or_mo_transform(C.byref(self.oracle.ob), C.byref(self.res), nao,
                                           _p(CA), CA.shape[0], a0, a1, _p(CB), CB.shape[0], b0, b1,
                                           max_buffer_mb, _p(out))
        if st != 0:
            raise RuntimeError("Buffer size too small for Rpq->Rab")
        return out.reshape(NA * NB, self.nvecs).T

    def __del__(self):
        try:
            self.oracle.L.or_chol_free(C.byref(self.res))
        except Exception:
            pass


def full_vectors(shell_pair_loc, shell_pair_dim, rkpq, norb_pairs):
    """(nvecs, norb_pairs) array R[k, f] = R_k at FULL location f + 1.
    shell_pair_loc[s] = (FULL, subset storage, subset index), 1-based, as in
    ShellPairLoc (TwoStepCholesky_definitions.f90:56-74); rkpq[X, i, k] is
    Rkpq(k, i + 1, X + 1)."""
    nvecs = rkpq.shape[2]
    R = np.zeros((nvecs, norb_pairs))
    for s in range(len(shell_pair_dim)):
        f0, i0, X = (int(v) for v in shell_pair_loc[s])
        d = int(shell_pair_dim[s])
        R[:, f0 - 1:f0 - 1 + d] = rkpq[X - 1, i0 - 1:i0 - 1 + d, :].T
    return R


def step2_columns(basis, meta):
    """0-based FULL locations of the orbital pairs Step 2 factorizes (the
    decoded pivots: PivotOrbPairs inside PivotShellPairs), in pivot order.
    The vectors reproduce exactly these columns of V."""
    sp = meta["shell_pairs"]
    loc = meta["shell_pair_loc"]
    psp = meta["pivot_shell_pairs"]
    psploc = meta["pivot_shell_pair_loc"]
    pspdim = meta["pivot_shell_pair_dim"]
    porb = meta["pivot_orb_pairs"]
    cols = []
    for k in range(len(psp)):
        ab = int(psp[k]) - 1
        A, B = sp[ab]
        na = basis.shell_nang(A - 1)
        f0 = int(loc[ab][0]) - 1
        for i in range(int(psploc[k]) - 1, int(psploc[k]) - 1 + int(pspdim[k])):
            fb = int(porb[i]) - 1
            a, b = fb % na, fb // na
            if A == B:
                if a < b:
                    a, b = b, a
                pos = sum(na - bb for bb in range(b)) + (a - b)
            else:
                pos = fb
            cols.append(f0 + pos)
    return np.array(cols, dtype=np.int64)


def boys_fm(m, x):
    F = np.zeros(m + 1)
    lib().or_boys_fm(m, float(x), _p(F))
    return F


def boys_fm_reference(m, x):
    return lib().or_boys_fm_reference(m, float(x))


def dsort0(x, y, kflag):
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    y = np.ascontiguousarray(y, dtype=np.int64).copy()
    lib().or_dsort0(_p(x), _p(y, i64), len(x), kflag)
    return x, y


def spher_u(l):
    nc = (l + 1) * (l + 2) // 2
    U = np.zeros((2 * l + 1) * nc)
    lib().or_spher_u(l, _p(U))
    return U.reshape(2 * l + 1, nc)


def pq2p_ge_q(pq, n):
    p, q = i64(), i64()
    lib().or_pq2p_ge_q(pq, n, C.byref(p), C.byref(q))
    return p.value, q.value
