"""External AO orderings on the device (gi_ao_map, gi_ao_rows_to_auto2e_device)
and through the Gammcor-facing mirror (chol_gammcor_Rkab,
ints1e_gammcor_H0_mo / _H0_extao):

* the library's maps and Orca signs equal the oracle's restatement of
  auto2e_interface_AOMap / ApplyOrcaPhases (exact integers);
* the device row transform equals the host application of the oracle map,
  bit for bit;
* coefficients written in any external order give exactly the Rkab and H0_mo
  of the same orbitals in the Auto2e order (the transforms are permutations
  and signs, so the products see identical inputs)."""
import numpy as np
import pytest

import aomap as A
from gigpu import gammcor, lib, oneelectron, systems

pytestmark = pytest.mark.gpu

SPHER_ORDERINGS = [A.MOLPRO, A.ORCA, A.DALTON, A.OPEN_MOLCAS, A.PYSCF]


@pytest.fixture(scope="module")
def spher_basis():
    return systems.water_system(2, 6.0, "large", seed=71, contracted=True)


@pytest.mark.parametrize("ordering", SPHER_ORDERINGS)
def test_library_map_equals_oracle(gpu_ctx, spher_basis, ordering):
    gpu_ctx.set_basis(spher_basis)
    m, ph = gpu_ctx.ao_map(ordering)
    ls, atoms = A.basis_shells(spher_basis)
    mo, pho = A.ao_map(ls, atoms, True, ordering)
    assert np.array_equal(m, mo)
    assert np.array_equal(ph, pho)


def test_cartesian_map_and_rejections(gpu_ctx, spher_basis):
    b = systems.water_system(1, 6.0, "large", seed=72, contracted=True, spher_ao=False)
    gpu_ctx.set_basis(b)
    m, ph = gpu_ctx.ao_map(A.MOLPRO)
    ls, atoms = A.basis_shells(b)
    mo, pho = A.ao_map(ls, atoms, False, A.MOLPRO)
    assert np.array_equal(m, mo) and np.array_equal(ph, pho)
    for o in (A.ORCA, A.DALTON, A.OPEN_MOLCAS, A.PYSCF, A.AUTO2E):
        with pytest.raises(lib.GigpuError, match="Invalid external orbital ordering"):
            gpu_ctx.ao_map(o)
    gpu_ctx.set_basis(spher_basis)
    with pytest.raises(lib.GigpuError, match="Invalid external orbital ordering"):
        gpu_ctx.ao_map(A.AUTO2E)


class _Hip:
    """Device buffers through the HIP runtime the library itself uses."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]

    def upload(self, arr):
        C = self.C
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), arr.nbytes) == 0
        assert self.L.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) == 0
        return p.value

    def download(self, ptr, out):
        C = self.C
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes, 2) == 0
        return out

    def free(self, ptr):
        self.L.hipFree(self.C.c_void_p(ptr))


def _to_external(C_ao, m, ph):
    """Inverse of the row transform: C_ao(m[p]-1) = ph(m[p]-1) C_ext(p)."""
    return np.asfortranarray(C_ao[m - 1] * ph[m - 1][:, None])


@pytest.mark.parametrize("ordering", [A.MOLPRO, A.ORCA, A.OPEN_MOLCAS])
def test_device_rows_and_rkab_invariance(gpu_ctx, spher_basis, ordering):
    b = spher_basis
    vecs = gammcor.chol_gammcor_Rkpq(b, gammcor.CHOL_ACCURACY_TIGHT, ctx=gpu_ctx)
    rng = np.random.default_rng(5)
    C_ao, _ = np.linalg.qr(rng.normal(size=(b.nao, b.nao)))
    C_ao = np.asfortranarray(C_ao)
    m, ph = gpu_ctx.ao_map(ordering)
    C_ext = _to_external(C_ao, m, ph)
    # device row transform, bit for bit against the host application
    hip = _Hip()
    X = hip.upload(np.asfortranarray(C_ext))                 # column-major nao x nmo
    Y = hip.upload(np.zeros_like(C_ext, order="F"))
    try:
        gpu_ctx.ao_rows_to_auto2e_device(ordering, X, b.nao, b.nao, Y, b.nao)
        got_rows = hip.download(Y, np.zeros_like(C_ext, order="F"))
    finally:
        hip.free(X)
        hip.free(Y)
    assert np.array_equal(got_rows, C_ao)
    # Rkab from external-order coefficients = Rkab from the Auto2e-order ones
    nocc = 8
    ref = gammcor.chol_gammcor_Rkab(None, C_ao, 1, nocc, C_ao, nocc + 1, nocc + 20, 100000, vecs, b,
                                    gammcor.ORBITAL_ORDERING_DALTON)
    got = gammcor.chol_gammcor_Rkab(None, C_ext, 1, nocc, C_ext, nocc + 1, nocc + 20, 100000, vecs, b, ordering)
    assert np.array_equal(got, ref)


def test_h0_mo_and_extao(gpu_ctx, spher_basis):
    b = spher_basis
    system = oneelectron.system_of(b)
    rng = np.random.default_rng(6)
    C_ao, _ = np.linalg.qr(rng.normal(size=(b.nao, b.nao)))
    H0_ao = oneelectron.ints1e_gammcor_H0(None, b, system, gpu_ctx)
    ref = oneelectron.ints1e_gammcor_H0_mo(None, C_ao, b, system, gammcor.ORBITAL_ORDERING_DALTON, gpu_ctx)
    for o in (A.MOLPRO, A.ORCA, A.OPEN_MOLCAS):
        m, ph = gpu_ctx.ao_map(o)
        got = oneelectron.ints1e_gammcor_H0_mo(None, _to_external(C_ao, m, ph), b, system, o, gpu_ctx)
        assert np.array_equal(got, ref)
        # H0_extao: both indices to the external order; for Orca the signs at
        # their Auto2e positions are then applied to the external-order matrix,
        # as OneElectronInts_Gammcor.f90:152-159 does
        ext = oneelectron.ints1e_gammcor_H0_extao(None, b, system, o, gpu_ctx)
        want = H0_ao[np.ix_(m - 1, m - 1)]
        if o == A.ORCA:
            want = want * ph[:, None] * ph[None, :]
        assert np.array_equal(ext, want)


def test_device_rows_error_paths(gpu_ctx, spher_basis):
    """In-place reordering and a leading dimension below NAO are rejected
    before any kernel runs."""
    b = spher_basis
    gpu_ctx.set_basis(b)
    hip = _Hip()
    X = hip.upload(np.zeros((b.nao, 2), order="F"))
    Y = hip.upload(np.zeros((b.nao, 2), order="F"))
    try:
        with pytest.raises(lib.GigpuError, match="in-place"):
            gpu_ctx.ao_rows_to_auto2e_device(A.MOLPRO, X, b.nao, 2, X, b.nao)
        with pytest.raises(lib.GigpuError, match="Invalid number of atomic orbitals"):
            gpu_ctx.ao_rows_to_auto2e_device(A.MOLPRO, X, b.nao - 1, 2, Y, b.nao)
        with pytest.raises(lib.GigpuError, match="Invalid external orbital ordering"):
            gpu_ctx.ao_rows_to_auto2e_device(A.AUTO2E, X, b.nao, 2, Y, b.nao)
    finally:
        hip.free(X)
        hip.free(Y)
