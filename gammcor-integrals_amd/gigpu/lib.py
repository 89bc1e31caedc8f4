"""ctypes binding of libgigpu (C ABI declared in include/gigpu.h).

The binding loads the in-tree ``gigpu/libgigpu.so``.  There is no fallback:
if the library is missing or no GPU is present, calls fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgigpu.so")

EXPORTS = ["gi_create", "gi_destroy", "gi_last_error", "gi_set_basis", "gi_eri_quartets",
           "gi_chol_build", "gi_chol_accuracy_tau", "gi_chol_dims", "gi_chol_meta", "gi_chol_rkpq",
           "gi_chol_rkpq_device", "gi_chol_inv_l", "gi_chol_vdiag", "gi_chol_timings",
           "gi_mo_transform", "gi_mo_transform_device", "gi_eri_class_stats", "gi_eri_pair_block",
           "gi_set_option", "gi_chol_build_batch", "gi_coulomb_energy_grad",
           "gi_chol_residual", "gi_one_electron", "gi_h0", "gi_ao_map", "gi_ao_rows_to_auto2e_device",
           "gi_ao_rows_to_auto2e_device_noalloc", "gi_thc_orbitals", "gi_thc_orbitals_device", "gi_thc_z",
           "gi_thc_z_device", "gi_thc_rkab", "gi_thc_rkab_device", "gi_thc_grid_pivots"]

CHOL_DECODE_FIXED = 1
CHOL_SKIP_VECTORS = 2
CHOL_TIME_ERI = 4
CHOL_RAW_TIES = 8
CHOL_SPLIT_BF16 = 16
MO_ALL_COLUMNS = 1
MO_SPLIT_BF16 = 2
ORDERING_AUTO2E, ORDERING_MOLPRO, ORDERING_ORCA, ORDERING_DALTON, ORDERING_OPEN_MOLCAS, ORDERING_PYSCF = range(6)

i64 = C.c_int64
dptr = C.POINTER(C.c_double)
iptr = C.POINTER(C.c_int64)

_lib = None


class GigpuError(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GigpuError("libgigpu.so is not built (%s); run `python -m gigpu.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    L.gi_last_error.restype = C.c_char_p
    L.gi_create.argtypes = [C.POINTER(C.c_void_p), i64]
    L.gi_destroy.argtypes = [C.c_void_p]
    L.gi_set_basis.argtypes = [C.c_void_p, i64, dptr, i64, iptr, iptr, i64, iptr, iptr, i64, dptr, dptr, i64,
                               dptr, i64]
    L.gi_eri_quartets.argtypes = [C.c_void_p, i64, iptr, C.c_double, iptr, dptr, i64]
    L.gi_chol_build.argtypes = [C.c_void_p, C.c_double, C.c_double, i64, i64, i64]
    L.gi_chol_accuracy_tau.argtypes = [i64, dptr]
    L.gi_chol_dims.argtypes = [C.c_void_p, iptr]
    L.gi_chol_meta.argtypes = [C.c_void_p] + [iptr] * 11
    L.gi_chol_rkpq.argtypes = [C.c_void_p, dptr]
    L.gi_chol_rkpq_device.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    L.gi_chol_inv_l.argtypes = [C.c_void_p, dptr]
    L.gi_chol_vdiag.argtypes = [C.c_void_p, dptr]
    L.gi_chol_timings.argtypes = [C.c_void_p, dptr]
    L.gi_eri_class_stats.argtypes = [C.c_void_p, i64, dptr, iptr]
    L.gi_eri_pair_block.argtypes = [C.c_void_p, i64, iptr, i64, iptr, C.c_double, dptr, i64]
    L.gi_set_option.argtypes = [C.c_void_p, C.c_char_p, i64]
    L.gi_chol_build_batch.argtypes = [i64, C.POINTER(C.c_void_p), C.c_double, C.c_double, i64, i64, i64, i64,
                                      iptr]
    L.gi_chol_residual.argtypes = [C.c_void_p, dptr, dptr]
    L.gi_one_electron.argtypes = [C.c_void_p, C.c_void_p, i64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, i64]
    L.gi_h0.argtypes = [C.c_void_p, C.c_void_p, i64, C.c_void_p, C.c_void_p, dptr, i64]
    L.gi_coulomb_energy_grad.argtypes = [C.c_void_p, dptr, i64, C.c_double, dptr, C.c_void_p]
    L.gi_mo_transform.argtypes = [C.c_void_p, dptr, i64, i64, i64, dptr, i64, i64, i64, i64, i64, dptr]
    L.gi_mo_transform_device.argtypes = [C.c_void_p, C.c_void_p, i64, i64, i64, C.c_void_p, i64, i64, i64, i64,
                                         i64, C.c_void_p]
    L.gi_ao_map.argtypes = [C.c_void_p, i64, iptr, dptr]
    L.gi_ao_rows_to_auto2e_device.argtypes = [C.c_void_p, i64, C.c_void_p, i64, i64, C.c_void_p, i64]
    L.gi_ao_rows_to_auto2e_device_noalloc.argtypes = [C.c_void_p, i64, C.c_void_p, i64, i64, C.c_void_p, i64]
    L.gi_thc_orbitals.argtypes = [C.c_void_p, i64, dptr, i64, dptr, i64]
    L.gi_thc_orbitals_device.argtypes = [C.c_void_p, i64, C.c_void_p, i64, C.c_void_p, i64]
    L.gi_thc_z.argtypes = [C.c_void_p, i64, dptr, i64, i64, dptr, i64, C.c_void_p, i64]
    L.gi_thc_z_device.argtypes = [C.c_void_p, i64, C.c_void_p, i64, i64, C.c_void_p, i64, C.c_void_p, i64]
    L.gi_thc_rkab.argtypes = [C.c_void_p, i64, dptr, i64, dptr, i64, i64, dptr, i64, i64, i64, dptr, i64, i64, i64,
                              i64, dptr]
    L.gi_thc_grid_pivots.argtypes = [C.c_void_p, i64, dptr, C.c_double, C.c_double, i64, iptr, iptr, iptr]
    L.gi_thc_rkab_device.argtypes = [C.c_void_p, i64, C.c_void_p, i64, C.c_void_p, i64, i64, C.c_void_p, i64, i64,
                                     i64, C.c_void_p, i64, i64, i64, i64, C.c_void_p]
    for name in EXPORTS:
        getattr(L, name).restype = C.c_int if name != "gi_last_error" else C.c_char_p
    _lib = L
    return L


def _check(status):
    if status != 0:
        msg = _lib.gi_last_error().decode(errors="replace")
        raise GigpuError("gigpu error %d: %s" % (status, msg))


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class Context:
    """One libgigpu context bound to one HIP device."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = C.c_void_p()
        _check(self.L.gi_create(C.byref(h), device))
        self.h = h
        self.basis = None

    def close(self):
        if getattr(self, "h", None):
            self.L.gi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- basis -----------------------------------------------------------
    def set_basis(self, basis):
        a = basis.c_arrays()
        self._basis_arrays = a
        _check(self.L.gi_set_basis(self.h, basis.natoms, _p(a["coords"]), basis.nshells,
                                   _p(a["shell_centers"], i64), _p(a["shell_params"], i64), basis.nparams,
                                   _p(a["momentum"], i64), _p(a["nprim"], i64), a["cntr"].shape[1],
                                   _p(a["cntr"]), _p(a["exps"]), a["norm"].shape[1], _p(a["norm"]),
                                   1 if basis.spher_ao else 0))
        self.basis = basis

    # -- ERIs ------------------------------------------------------------
    def eri_quartets(self, quartets, kappa=0.0):
        """List of V[a,b,c,d] blocks for 0-based shell quartets."""
        q = np.ascontiguousarray(np.asarray(quartets, dtype=np.int64).reshape(-1, 4))
        dims = [[self.basis.shell_nang(int(s)) for s in row] for row in q]
        sizes = np.array([d[0] * d[1] * d[2] * d[3] for d in dims], dtype=np.int64)
        offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        out = np.zeros(int(sizes.sum()) if len(sizes) else 1)
        _check(self.L.gi_eri_quartets(self.h, len(q), _p(q, i64), kappa, _p(offs, i64), _p(out), len(out)))
        blocks = []
        for (na, nb, nc, nd), o, n in zip(dims, offs, sizes):
            blocks.append(out[o:o + n].reshape(nd, nc, nb, na).transpose(3, 2, 1, 0))
        return blocks

    def eri_pair_block(self, row_pairs, col_pairs, kappa=0.0):
        """Dense (pq|rs) block: rows = orbital pairs of the row shell pairs
        (0-based (A, B), a fastest), columns likewise."""
        rp = np.ascontiguousarray(np.asarray(row_pairs, dtype=np.int64).reshape(-1, 2))
        cp = np.ascontiguousarray(np.asarray(col_pairs, dtype=np.int64).reshape(-1, 2))
        nr = sum(self.basis.shell_nang(int(a)) * self.basis.shell_nang(int(b)) for a, b in rp)
        nc = sum(self.basis.shell_nang(int(a)) * self.basis.shell_nang(int(b)) for a, b in cp)
        out = np.zeros(nr * nc)
        _check(self.L.gi_eri_pair_block(self.h, len(rp), _p(rp, i64), len(cp), _p(cp, i64), kappa, _p(out), nr))
        return out.reshape(nc, nr).T

    def set_option(self, name, value):
        _check(self.L.gi_set_option(self.h, name.encode(), int(value)))

    # -- Cholesky --------------------------------------------------------
    def chol_build(self, tau=1e-6, kappa=0.0, max_block_dim=4000, vec_block=500, flags=0):
        _check(self.L.gi_chol_build(self.h, tau, kappa, max_block_dim, vec_block, flags))

    def chol_dims(self):
        d = np.zeros(8, np.int64)
        _check(self.L.gi_chol_dims(self.h, _p(d, i64)))
        return dict(nvecs=int(d[0]), norb_pairs=int(d[1]), nshell_pairs=int(d[2]), nsubsets=int(d[3]),
                    max_subset_dim=int(d[4]), npivot_shell_pairs=int(d[5]), nmacro=int(d[6]),
                    nquartets=int(d[7]))

    def chol_meta(self):
        d = self.chol_dims()
        nsp, nsub, nop, nv, npsp = (d["nshell_pairs"], d["nsubsets"], d["norb_pairs"], d["nvecs"],
                                    d["npivot_shell_pairs"])
        arrs = dict(shell_pairs=np.zeros(2 * nsp, np.int64), shell_pair_loc=np.zeros(3 * nsp, np.int64),
                    shell_pair_dim=np.zeros(nsp, np.int64), subset_dim=np.zeros(nsub, np.int64),
                    subset_bounds=np.zeros(2 * nsub, np.int64), pivots=np.zeros(nop, np.int64),
                    pivot_order=np.zeros(max(nv, 1), np.int64), pivot_shell_pairs=np.zeros(max(npsp, 1), np.int64),
                    pivot_shell_pair_loc=np.zeros(max(npsp, 1), np.int64),
                    pivot_shell_pair_dim=np.zeros(max(npsp, 1), np.int64),
                    pivot_orb_pairs=np.zeros(max(nv, 1), np.int64))
        keys = ["shell_pairs", "shell_pair_loc", "shell_pair_dim", "subset_dim", "subset_bounds", "pivots",
                "pivot_order", "pivot_shell_pairs", "pivot_shell_pair_loc", "pivot_shell_pair_dim",
                "pivot_orb_pairs"]
        _check(self.L.gi_chol_meta(self.h, *[_p(arrs[k], i64) for k in keys]))
        arrs["shell_pairs"] = arrs["shell_pairs"].reshape(nsp, 2)
        arrs["shell_pair_loc"] = arrs["shell_pair_loc"].reshape(nsp, 3)
        arrs["subset_bounds"] = arrs["subset_bounds"].reshape(nsub, 2)
        arrs["pivot_order"] = arrs["pivot_order"][:nv]
        arrs["pivot_orb_pairs"] = arrs["pivot_orb_pairs"][:nv]
        for k in ("pivot_shell_pairs", "pivot_shell_pair_loc", "pivot_shell_pair_dim"):
            arrs[k] = arrs[k][:npsp]
        arrs.update(d)
        return arrs

    def chol_rkpq(self):
        d = self.chol_dims()
        out = np.zeros(d["nvecs"] * d["max_subset_dim"] * d["nsubsets"])
        _check(self.L.gi_chol_rkpq(self.h, _p(out)))
        return out.reshape(d["nsubsets"], d["max_subset_dim"], d["nvecs"])

    def chol_rkpq_device_ptr(self):
        p = C.c_void_p()
        _check(self.L.gi_chol_rkpq_device(self.h, C.byref(p)))
        return p.value

    def chol_inv_l(self):
        n = self.chol_dims()["nvecs"]
        out = np.zeros(n * n)
        _check(self.L.gi_chol_inv_l(self.h, _p(out)))
        return out.reshape(n, n).T

    @staticmethod
    def _charges(natoms, charges, point_charges):
        q = None if charges is None else np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
        if q is not None and q.shape != (natoms,):
            raise ValueError("one charge per atom")
        pc = np.zeros((0, 4)) if point_charges is None else \
            np.ascontiguousarray(np.asarray(point_charges, dtype=np.float64).reshape(-1, 4))
        xyz = np.ascontiguousarray(pc[:, :3])
        qq = np.ascontiguousarray(pc[:, 3])
        return q, len(pc), xyz, qq

    def one_electron(self, charges, point_charges=None, ld=None):
        """(S, T, V) in the AO basis: gi_one_electron.  charges: one per atom;
        point_charges: rows (x, y, z, q), bohr; ld: leading dimension of the
        output arrays (>= nao; the returned views are nao x nao)."""
        n = self.basis.nao
        ld = n if ld is None else ld
        q, npc, xyz, qq = self._charges(self.basis.natoms, charges, point_charges)
        S, T, V = (np.zeros((ld, n), order="F") for _ in range(3))
        _check(self.L.gi_one_electron(self.h, None if q is None else q.ctypes.data, npc,
                                      xyz.ctypes.data if npc else None, qq.ctypes.data if npc else None,
                                      S.ctypes.data, T.ctypes.data, V.ctypes.data, ld))
        return S[:n], T[:n], V[:n]

    def h0(self, charges, point_charges=None, ld=None):
        """H0 = T + V_ne in the AO basis: gi_h0 (ints1e_gammcor_H0); ld as in
        one_electron."""
        n = self.basis.nao
        ld = n if ld is None else ld
        q, npc, xyz, qq = self._charges(self.basis.natoms, charges, point_charges)
        H = np.zeros((ld, n), order="F")
        _check(self.L.gi_h0(self.h, None if q is None else q.ctypes.data, npc, xyz.ctypes.data if npc else None,
                            qq.ctypes.data if npc else None, _p(H), ld))
        return H[:n]

    def chol_residual(self):
        """(max, min) over BASE orbital pairs of (pq|pq) - sum_k Rkpq(k, pq)^2."""
        mx, mn = np.zeros(1), np.zeros(1)
        _check(self.L.gi_chol_residual(self.h, _p(mx), _p(mn)))
        return float(mx[0]), float(mn[0])

    def coulomb_energy(self, density, kappa=0.0, grad=True):
        """E_J = 1/2 sum D_pq (pq|rs) D_rs of a symmetric AO density (nao x nao),
        and with grad=True also dE_J/dR as an (natoms, 3) array (Eh/bohr):
        gi_coulomb_energy_grad.  Returns E or (E, G)."""
        n = self.basis.nao
        D = np.asfortranarray(np.asarray(density, dtype=np.float64))
        if D.shape != (n, n):
            raise ValueError("density must be nao x nao (%d)" % n)
        e = np.zeros(1)
        if grad:
            g = np.zeros((self.basis.natoms, 3))
            _check(self.L.gi_coulomb_energy_grad(self.h, _p(D), n, float(kappa), _p(e), g.ctypes.data))
            return float(e[0]), g
        _check(self.L.gi_coulomb_energy_grad(self.h, _p(D), n, float(kappa), _p(e), None))
        return float(e[0])

    def chol_vdiag(self):
        n = self.basis.nao
        out = np.zeros(n * (n + 1) // 2)
        _check(self.L.gi_chol_vdiag(self.h, _p(out)))
        return out

    def chol_timings(self):
        t = np.zeros(18)
        _check(self.L.gi_chol_timings(self.h, _p(t)))
        keys = ["step1", "step2", "vectors", "total", "eri_kernel", "eri_flops", "vdiag", "eri_step1", "rtr",
                "micro", "vpqrs", "potrf_trtri", "vabrs", "vectors_gemm", "eri_launches", "rtr_flops",
                "candidates", "batch"]
        return dict(zip(keys, t.tolist()))

    def eri_class_stats(self):
        """[(E, F, quartets, kernel_ms, fp64_ops)] per kernel (la+lb, lc+ld) (needs GIGPU_ERI_STATS)."""
        rows = np.zeros(5 * 512)
        n = C.c_int64()
        _check(self.L.gi_eri_class_stats(self.h, 512, _p(rows), C.byref(n)))
        return [tuple(rows[5 * k:5 * k + 5]) for k in range(n.value)]

    def mo_transform(self, CA, a0, a1, CB, b0, b1, max_buffer_mb=1 << 20, flags=0):
        CA = np.asfortranarray(CA, dtype=np.float64)
        CB = np.asfortranarray(CB, dtype=np.float64)
        nv = self.chol_dims()["nvecs"]
        NA, NB = a1 - a0 + 1, b1 - b0 + 1
        out = np.zeros(nv * NA * NB)
        _check(self.L.gi_mo_transform(self.h, _p(CA), CA.shape[0], a0, a1, _p(CB), CB.shape[0], b0, b1,
                                      max_buffer_mb, flags, _p(out)))
        return out.reshape(NA * NB, nv).T

    def mo_transform_device(self, CA_ptr, lda, a0, a1, CB_ptr, ldb, b0, b1, Rkab_ptr, max_buffer_mb=1 << 20,
                            flags=0):
        _check(self.L.gi_mo_transform_device(self.h, C.c_void_p(CA_ptr), lda, a0, a1, C.c_void_p(CB_ptr), ldb,
                                             b0, b1, max_buffer_mb, flags, C.c_void_p(Rkab_ptr)))

    def ao_map(self, ordering):
        """(map, phase): map[p] (1-based) = Auto2e position of external AO p,
        phase in Auto2e order (Orca signs), for the context's basis."""
        n = self.basis.nao
        m = np.zeros(n, dtype=np.int64)
        ph = np.zeros(n)
        _check(self.L.gi_ao_map(self.h, ordering, _p(m, i64), _p(ph)))
        return m, ph

    def ao_rows_to_auto2e_device(self, ordering, X_ptr, ld_ext, ncols, Y_ptr, ld_ao):
        _check(self.L.gi_ao_rows_to_auto2e_device(self.h, ordering, C.c_void_p(X_ptr), ld_ext, ncols,
                                                  C.c_void_p(Y_ptr), ld_ao))

    def ao_rows_to_auto2e_device_noalloc(self, ordering, X_ptr, ld_ext, ncols, Y_ptr, ld_ao):
        """gi_ao_rows_to_auto2e_device_noalloc: the same transform without device
        allocations (map and signs staged in the library's static array)."""
        _check(self.L.gi_ao_rows_to_auto2e_device_noalloc(self.h, ordering, C.c_void_p(X_ptr), ld_ext, ncols,
                                                          C.c_void_p(Y_ptr), ld_ao))

    def thc_grid_pivots(self, xyz, qr_thresh, qr_thresh_reduced=-1.0, block_dim=500):
        """gi_thc_grid_pivots: (pivots 0-based in pivot order, npivots_reduced)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        n = xyz.shape[0]
        piv = np.zeros(max(n, 1), np.int64)
        npv, nred = i64(), i64()
        _check(self.L.gi_thc_grid_pivots(self.h, n, _p(xyz), qr_thresh, qr_thresh_reduced, block_dim,
                                         C.byref(npv), C.byref(nred), _p(piv, i64)))
        return piv[:npv.value].copy(), nred.value

    def thc_orbitals(self, xyz, normalize=True):
        """gi_thc_orbitals: Xgp (npts, nao) of the points xyz (npts, 3), bohr."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        n = xyz.shape[0]
        X = np.zeros((n, self.basis.nao), order="F")
        _check(self.L.gi_thc_orbitals(self.h, n, _p(xyz), int(bool(normalize)), _p(X), max(n, 1)))
        return X

    def thc_orbitals_device(self, npts, xyz_ptr, X_ptr, ldx, normalize=True):
        _check(self.L.gi_thc_orbitals_device(self.h, npts, C.c_void_p(xyz_ptr), int(bool(normalize)),
                                             C.c_void_p(X_ptr), ldx))

    def thc_z(self, Xgp, npts_reduced=0):
        """gi_thc_z: (Zgk, Zgk_reduced or None) for the context's Cholesky vectors."""
        X = np.asfortranarray(Xgp, dtype=np.float64)
        n = X.shape[0]
        nv = self.chol_dims()["nvecs"]
        Z = np.zeros((n, nv), order="F")
        red = 0 < npts_reduced < n
        Zr = np.zeros((npts_reduced, nv), order="F") if red else None
        _check(self.L.gi_thc_z(self.h, n, _p(X), n, npts_reduced, _p(Z), n,
                               Zr.ctypes.data_as(C.c_void_p) if red else None, npts_reduced if red else 1))
        return Z, Zr

    def thc_z_device(self, npts, X_ptr, ldx, Z_ptr, ldz, npts_reduced=0, Zr_ptr=0, ldzr=1):
        _check(self.L.gi_thc_z_device(self.h, npts, C.c_void_p(X_ptr), ldx, npts_reduced, C.c_void_p(Z_ptr), ldz,
                                      C.c_void_p(Zr_ptr) if Zr_ptr else None, ldzr))

    def thc_rkab(self, Xgp, Zgk, CA, a0, a1, CB, b0, b1, ordering=0):
        """gi_thc_rkab: Rkab (nvec, NA*NB), a fastest."""
        X = np.asfortranarray(Xgp, dtype=np.float64)
        Z = np.asfortranarray(Zgk, dtype=np.float64)
        CA = np.asfortranarray(CA, dtype=np.float64)
        CB = np.asfortranarray(CB, dtype=np.float64)
        n, nv = Z.shape
        NA, NB = a1 - a0 + 1, b1 - b0 + 1
        out = np.zeros(nv * NA * NB)
        _check(self.L.gi_thc_rkab(self.h, n, _p(X), X.shape[0], _p(Z), n, nv, _p(CA), CA.shape[0], a0, a1, _p(CB),
                                  CB.shape[0], b0, b1, ordering, _p(out)))
        return out.reshape(NA * NB, nv).T

    def thc_rkab_device(self, npts, X_ptr, ldx, Z_ptr, ldz, nvec, CA_ptr, lda, a0, a1, CB_ptr, ldb, b0, b1,
                        Rkab_ptr, ordering=0):
        _check(self.L.gi_thc_rkab_device(self.h, npts, C.c_void_p(X_ptr), ldx, C.c_void_p(Z_ptr), ldz, nvec,
                                         C.c_void_p(CA_ptr), lda, a0, a1, C.c_void_p(CB_ptr), ldb, b0, b1, ordering,
                                         C.c_void_p(Rkab_ptr)))


def accuracy_tau(accuracy: int) -> float:
    L = load()
    t = C.c_double()
    _check(L.gi_chol_accuracy_tau(accuracy, C.byref(t)))
    return t.value


def chol_build_batch(contexts, tau=1e-6, kappa=0.0, max_block_dim=4000, vec_block=500, flags=0, nworkers=8):
    """Decompose the systems of several contexts (each with its basis set)
    concurrently: gi_chol_build_batch.  Returns the per-context statuses."""
    L = load()
    n = len(contexts)
    handles = (C.c_void_p * max(n, 1))(*[c.h.value for c in contexts])
    status = np.zeros(max(n, 1), np.int64)
    _check(L.gi_chol_build_batch(n, handles, tau, kappa, max_block_dim, vec_block, flags, nworkers,
                                 _p(status, i64)))
    return status[:n]
