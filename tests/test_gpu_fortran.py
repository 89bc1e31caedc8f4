"""The Fortran side of the boundary on the GPU: tests/fortran/chol_check calls
gigpu_chol_Rkpq and gigpu_chol_Rkab (module gigpu_cholesky, the wrappers of
chol_gammcor_Rkpq / chol_gammcor_Rkab, Cholesky_Gammcor.f90:26-92,335-377)
through ISO_C_BINDING on a small basis; the TChol2Vecs dimensions and index
arrays, Inv_L, Rkpq and Rkab it returns equal those of the ctypes path on
the same basis, and the pivots equal the oracle's."""
import os
import subprocess

import numpy as np
import pytest

from gigpu import systems

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "build", "chol_check")


def _write_input(path, b, accuracy, CA, CB, a0, a1, b0, b1, max_buffer_mb):
    a = b.c_arrays()
    head = np.array([b.natoms, b.nshells, b.nparams, a["cntr"].shape[1], a["norm"].shape[1],
                     1 if b.spher_ao else 0, accuracy, CA.shape[0], CA.shape[1], a0, a1, b0, b1, max_buffer_mb],
                    np.int64)
    with open(path, "wb") as f:
        f.write(head.tobytes())
        # C-contiguous (n, m) arrays are the Fortran (m, n) column-major arrays
        for k in ("coords",):
            f.write(a[k].tobytes())
        for k in ("shell_centers", "shell_params", "momentum", "nprim"):
            f.write(a[k].astype(np.int64).tobytes())
        for k in ("cntr", "exps", "norm"):
            f.write(a[k].tobytes())
        f.write(np.asfortranarray(CA).tobytes(order="F"))
        f.write(np.asfortranarray(CB).tobytes(order="F"))


def _read_output(path, na, nb):
    raw = open(path, "rb").read()
    i = np.frombuffer(raw, np.int64)
    nv, nop, nsp, nsub, msd, npsp = (int(x) for x in i[:6])
    o = 6
    out = dict(nvecs=nv, norb_pairs=nop, nshell_pairs=nsp, nsubsets=nsub, max_subset_dim=msd,
               npivot_shell_pairs=npsp)
    for k, n in (("shell_pairs", 2 * nsp), ("shell_pair_loc", 3 * nsp), ("pivots", nop), ("pivot_orb_pairs", nv)):
        out[k] = i[o:o + n].copy()
        o += n
    d = np.frombuffer(raw[8 * o:], np.float64)
    out["inv_l"] = d[:nv * nv].reshape(nv, nv).T
    d = d[nv * nv:]
    out["rkpq"] = d[:nv * msd * nsub].reshape(nsub, msd, nv)
    d = d[nv * msd * nsub:]
    out["rkab"] = d[:nv * na * nb].reshape(na * nb, nv).T
    assert len(d) == nv * na * nb
    return out


@pytest.mark.parametrize("contracted", [False, True], ids=["uncontracted", "contracted"])
def test_fortran_wrappers_equal_ctypes_path(gpu_ctx, oracle_mod, tmp_path, contracted):
    if not os.path.exists(EXE):
        pytest.fail("tests/fortran/build/chol_check is not built (run __graft_entry__.build())")
    b = systems.water_system(2, 6.0, "small", seed=21, contracted=contracted)
    rng = np.random.default_rng(4)
    C = np.linalg.qr(rng.normal(size=(b.nao, b.nao)))[0]
    a0, a1, b0, b1 = 2, 6, 5, 17   # a0 > 1: the reference's column mask applies
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, b, 2, C, C, a0, a1, b0, b1, 1 << 20)
    r = subprocess.run([EXE, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fortran chol ok" in r.stdout
    f = _read_output(fout, a1 - a0 + 1, b1 - b0 + 1)

    gpu_ctx.set_basis(b)
    gpu_ctx.chol_build(tau=1e-6)
    d = gpu_ctx.chol_dims()
    m = gpu_ctx.chol_meta()
    for k in ("nvecs", "norb_pairs", "nshell_pairs", "nsubsets", "max_subset_dim", "npivot_shell_pairs"):
        assert f[k] == d[k], k
    for k in ("shell_pairs", "shell_pair_loc", "pivots", "pivot_orb_pairs"):
        assert np.array_equal(f[k], np.asarray(m[k]).ravel()[:len(f[k])]), k
    assert np.array_equal(f["inv_l"], gpu_ctx.chol_inv_l())
    assert np.array_equal(f["rkpq"], gpu_ctx.chol_rkpq())
    want = gpu_ctx.mo_transform(C, a0, a1, C, b0, b1)
    assert np.abs(f["rkab"] - want).max() <= 1e-14 * np.abs(want).max()
    ref = oracle_mod.Oracle(b).chol(tau=1e-6)
    assert ref.status == 0, ref.message
    assert f["nvecs"] == ref.nvecs
    assert np.array_equal(m["pivot_order"][:ref.nvecs], ref.pivot_order)


THC_EXE = os.path.join(ROOT, "tests", "fortran", "build", "thc_check")


def test_fortran_thc_equals_ctypes_path(gpu_ctx, tmp_path):
    """tests/fortran/thc_check drives gi_thc_grid_pivots, gi_chol_build,
    gi_thc_orbitals, gi_thc_z (with a reduced grid) and gi_thc_rkab (Orca
    ordering) from Fortran; pivots, Xgp, Zgk, ZgkReduced and Rkab equal the
    ctypes path on the same inputs, and the pivots the oracle's."""
    import thc
    if not os.path.exists(THC_EXE):
        pytest.fail("tests/fortran/build/thc_check is not built (run __graft_entry__.build())")
    b = systems.water_system(1, 6.0, "small", seed=1)
    parent = systems.atom_centred_points(b, radii=(0.3, 0.8, 1.5, 2.5, 4.0), ndir=24, seed=6)
    rng = np.random.default_rng(9)
    C = rng.standard_normal((b.nao, 10))
    a0, a1, b0, b1, ordering, block = 1, 6, 3, 10, 2, 50
    tau, kappa, qr, qr_red = 1e-6, 0.0, 1e-3, 3e-3
    a = b.c_arrays()
    head = np.array([b.natoms, b.nshells, b.nparams, a["cntr"].shape[1], a["norm"].shape[1], 1 if b.spher_ao else 0,
                     len(parent), block, C.shape[1], a0, a1, b0, b1, ordering], np.int64)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        f.write(head.tobytes())
        f.write(np.array([tau, kappa, qr, qr_red]).tobytes())
        f.write(a["coords"].tobytes())
        for k in ("shell_centers", "shell_params", "momentum", "nprim"):
            f.write(a[k].astype(np.int64).tobytes())
        for k in ("cntr", "exps", "norm"):
            f.write(a[k].tobytes())
        f.write(np.ascontiguousarray(parent).tobytes())
        f.write(np.asfortranarray(C).tobytes(order="F"))
        f.write(np.asfortranarray(C).tobytes(order="F"))
    r = subprocess.run([THC_EXE, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fortran thc ok" in r.stdout
    raw = open(fout, "rb").read()
    npiv, nred, nv = (int(x) for x in np.frombuffer(raw[:24], np.int64))
    piv = np.frombuffer(raw[24:24 + 8 * npiv], np.int64)
    d = np.frombuffer(raw[24 + 8 * npiv:], np.float64)
    X = d[:npiv * b.nao].reshape(b.nao, npiv).T
    d = d[npiv * b.nao:]
    Z = d[:npiv * nv].reshape(nv, npiv).T
    d = d[npiv * nv:]
    assert 0 < nred < npiv
    Zr = d[:nred * nv].reshape(nv, nred).T
    d = d[nred * nv:]
    NA, NB = a1 - a0 + 1, b1 - b0 + 1
    assert len(d) == nv * NA * NB
    Rk = d.reshape(NA * NB, nv).T

    gpu_ctx.set_basis(b)
    gpiv, gnred = gpu_ctx.thc_grid_pivots(parent, qr, qr_red, block)
    assert np.array_equal(piv, gpiv) and nred == gnred
    opiv, onred = thc.grid_pivots(b, parent, qr, qr_red, block)
    assert np.array_equal(piv, opiv) and nred == onred
    gX = gpu_ctx.thc_orbitals(parent[gpiv])
    assert np.array_equal(X, gX)
    gpu_ctx.chol_build(tau=tau)
    gZ, gZr = gpu_ctx.thc_z(gX, nred)
    assert np.array_equal(Z, gZ) and np.array_equal(Zr, gZr)
    gR = gpu_ctx.thc_rkab(gX, gZ, C, a0, a1, C, b0, b1, ordering)
    assert np.array_equal(Rk, gR)
