"""gi_ao_rows_to_auto2e_device_noalloc (the allocation-free variant of the
device AO row transform) against gi_ao_rows_to_auto2e_device: bit-identical
rows on two calls in a row, the second with a smaller basis and fewer
columns, through the C ABI; AO counts beyond the static staging array are
rejected."""
import numpy as np
import pytest

import aomap as A
from gigpu import lib, systems

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers through the HIP runtime the library itself uses."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def upload(self, arr):
        C = self.C
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), arr.nbytes) == 0
        assert self.L.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) == 0
        self.ptrs.append(p.value)
        return p.value

    def download(self, ptr, out):
        assert self.L.hipMemcpy(out.ctypes.data_as(self.C.c_void_p), self.C.c_void_p(ptr), out.nbytes, 2) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.L.hipFree(self.C.c_void_p(p))


def _rows(ctx, basis, ordering, ncols, ld_ao, seed, noalloc):
    hip = _Hip()
    try:
        rng = np.random.default_rng(seed)
        X = np.asfortranarray(rng.normal(size=(basis.nao, ncols)))
        Y0 = np.asfortranarray(rng.normal(size=(ld_ao, ncols)))   # padding rows must stay as they are
        dX, dY = hip.upload(X), hip.upload(Y0)
        f = ctx.ao_rows_to_auto2e_device_noalloc if noalloc else ctx.ao_rows_to_auto2e_device
        f(ordering, dX, basis.nao, ncols, dY, ld_ao)
        return hip.download(dY, np.zeros_like(Y0, order="F"))
    finally:
        hip.close()


@pytest.mark.parametrize("ordering", [A.ORCA, A.OPEN_MOLCAS, A.MOLPRO])
def test_noalloc_rows_equal_allocating_entry_point(ordering):
    ctx = lib.Context(0)
    try:
        big = systems.water_system(2, 6.0, "large", seed=81, contracted=True)
        small = systems.water_system(1, 6.0, "small", seed=82)
        assert small.nao < big.nao
        for basis, ncols, pad, seed in ((big, big.nao, 3, 1), (small, 7, 0, 2)):
            ctx.set_basis(basis)
            want = _rows(ctx, basis, ordering, ncols, basis.nao + pad, seed, noalloc=False)
            got = _rows(ctx, basis, ordering, ncols, basis.nao + pad, seed, noalloc=True)
            assert np.array_equal(got, want)
    finally:
        ctx.close()


def test_noalloc_error_paths():
    ctx = lib.Context(0)
    hip = _Hip()
    try:
        b = systems.water_system(1, 6.0, "small", seed=83)
        ctx.set_basis(b)
        X = hip.upload(np.zeros((b.nao, 2), order="F"))
        Y = hip.upload(np.zeros((b.nao, 2), order="F"))
        with pytest.raises(lib.GigpuError, match="in-place"):
            ctx.ao_rows_to_auto2e_device_noalloc(A.MOLPRO, X, b.nao, 2, X, b.nao)
        with pytest.raises(lib.GigpuError, match="Invalid number of atomic orbitals"):
            ctx.ao_rows_to_auto2e_device_noalloc(A.MOLPRO, X, b.nao - 1, 2, Y, b.nao)
        with pytest.raises(lib.GigpuError, match="Invalid external orbital ordering"):
            ctx.ao_rows_to_auto2e_device_noalloc(A.AUTO2E, X, b.nao, 2, Y, b.nao)
    finally:
        hip.close()
        ctx.close()
