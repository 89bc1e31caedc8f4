"""Argument handling and buffer sizes of the C ABI on the GPU:

* gi_h0 / gi_one_electron with a leading dimension larger than nao;
* the decomposition at CHOL_ACCURACY_DEBUG (tau = 1e-10): the residual
  diagonal ends below tau, with the pivots of the oracle (the quantized
  comparison keys order near-ties only; thresholds compare raw values); with
  raw comparison keys, consistent with the oracle's raw-key build;
* gi_mo_transform in several passes when MaxBufferDimMB is small (the
  reference's buffer bound, Cholesky_Gammcor.f90:176-186) equals one pass;
* gi_chol_build_batch rejects a non-positive or NaN tau like gi_chol_build;
* the TChol2Vecs accessors fail once the basis changes (no stale data)."""
import math

import numpy as np
import pytest

from gigpu import lib, systems
from grad_helpers import tiny_basis
from test_gpu_configs import raw_tie_consistent

pytestmark = pytest.mark.gpu


def test_h0_with_padded_leading_dimension(gpu_ctx):
    basis = tiny_basis(lmax=3, contracted=True)
    gpu_ctx.set_basis(basis)
    q = np.array([8.0, 1.0, 1.0])
    n = basis.nao
    S, T, V = gpu_ctx.one_electron(q)
    for ld in (n, n + 3, 2 * n + 1):
        H = gpu_ctx.h0(q, ld=ld)
        assert H.shape == (n, n)
        assert np.abs(H - (T + V)).max() <= 1e-13 * np.abs(T + V).max()
        S2, T2, V2 = gpu_ctx.one_electron(q, ld=ld)
        assert np.array_equal(S2, S) and np.array_equal(T2, T) and np.array_equal(V2, V)


@pytest.mark.parametrize("flags", [1, 1 | 8])
def test_debug_accuracy_residual_below_tau(gpu_ctx, oracle_mod, flags):
    b = systems.water_system(2, 6.0, "small", seed=1)
    tau = lib.accuracy_tau(4)
    assert tau == 1e-10
    ref = oracle_mod.Oracle(b).chol(tau=tau, flags=(1 | 4) if flags & 8 else 1)
    assert ref.status == 0, ref.message
    assert ref.max_residual <= tau
    gpu_ctx.set_basis(b)
    gpu_ctx.chol_build(tau=tau, flags=flags)
    m = gpu_ctx.chol_meta()
    if flags & lib.CHOL_RAW_TIES:
        # raw keys: mathematically equal keys are ordered by last-bit
        # differences between the two implementations' ERIs
        raw_tie_consistent(gpu_ctx, oracle_mod.Oracle(b), ref, m, tau, fixed_decode=True)
    else:
        assert m["nvecs"] == ref.nvecs
        assert np.array_equal(m["pivot_order"], ref.pivot_order)
    rmax, rmin = gpu_ctx.chol_residual()
    assert rmax <= tau * (1 + 1e-6)
    # (pq|pq) - sum_k R_k(pq)^2 >= 0 in exact arithmetic; the vectors pass
    # through Inv_L, whose condition grows as tau falls (about 1e4 here), and
    # the recomputed residual carries rounding errors of about 1e-11 (measured
    # -1.2e-11).  Wrong vectors would leave residuals of the size of the
    # diagonal (about 1), so half of tau still separates round-off from a defect.
    assert rmin >= -0.5 * tau


def test_mo_transform_passes_equal_one_pass(gpu_ctx):
    b = systems.water_system(2, 6.0, "small", seed=9)
    gpu_ctx.set_basis(b)
    gpu_ctx.chol_build(tau=1e-6)
    nv = gpu_ctx.chol_dims()["nvecs"]
    n = b.nao
    rng = np.random.default_rng(4)
    C, _ = np.linalg.qr(rng.normal(size=(n, n)))
    C = np.asfortranarray(C)
    a0, a1, b0, b1 = 1, 23, 24, 60
    one = gpu_ctx.mo_transform(C, a0, a1, C, b0, b1, flags=1)
    # buffer for 5 a-columns of W(NCholesky, Nk, NAO): five passes
    mb = int(math.ceil(5 * nv * n * 8 / 2 ** 20))
    maxnk = (mb * 2 ** 20 // 8) // (nv * n)
    assert 1 <= maxnk < a1 - a0 + 1
    many = gpu_ctx.mo_transform(C, a0, a1, C, b0, b1, max_buffer_mb=mb, flags=1)
    assert np.abs(many - one).max() <= 1e-13 * np.abs(one).max()
    # the reference's pass bounds: same columns as one pass (a0 = 1 skips none)
    ref_passes = gpu_ctx.mo_transform(C, a0, a1, C, b0, b1, max_buffer_mb=mb, flags=0)
    assert np.abs(ref_passes - one).max() <= 1e-13 * np.abs(one).max()
    # a0 > 1: the reference's passes skip some a-columns, which stay zero;
    # the columns they cover equal the single pass
    a0, a1 = 4, 23
    full = gpu_ctx.mo_transform(C, a0, a1, C, b0, b1, flags=1).reshape(nv, a1 - a0 + 1, b1 - b0 + 1, order="F")
    quirk = gpu_ctx.mo_transform(C, a0, a1, C, b0, b1, max_buffer_mb=mb, flags=0)
    quirk = quirk.reshape(nv, a1 - a0 + 1, b1 - b0 + 1, order="F")
    NA = a1 - a0 + 1
    npass = NA // maxnk + (NA % maxnk > 0)
    done = np.zeros(NA, bool)
    for k in range(1, npass + 1):
        lo, hi = (k - 1) * maxnk + a0, min(k * maxnk, a1)
        done[lo - a0:hi - a0 + 1] = True
    assert (~done).any()
    assert np.abs(quirk[:, done, :] - full[:, done, :]).max() <= 1e-13 * np.abs(full).max()
    assert not quirk[:, ~done, :].any()


@pytest.mark.parametrize("tau", [0.0, -1e-6, float("nan")])
def test_batch_rejects_invalid_tau(tau):
    c = lib.Context(0)
    try:
        c.set_basis(systems.water_system(1, 6.0, "small", seed=2))
        with pytest.raises(lib.GigpuError, match="tau must be positive"):
            lib.chol_build_batch([c], tau=tau)
        with pytest.raises(lib.GigpuError, match="MaxBlockDim"):
            lib.chol_build_batch([c], tau=1e-6, max_block_dim=0)
    finally:
        c.close()


def test_accessors_fail_after_basis_change(gpu_ctx):
    b = systems.water_system(1, 6.0, "small", seed=3)
    gpu_ctx.set_basis(b)
    gpu_ctx.chol_build(tau=1e-6)
    assert gpu_ctx.chol_dims()["nvecs"] > 0
    gpu_ctx.set_basis(b)
    for call in (gpu_ctx.chol_dims, gpu_ctx.chol_meta, gpu_ctx.chol_inv_l, gpu_ctx.chol_vdiag, gpu_ctx.chol_rkpq):
        with pytest.raises(lib.GigpuError):
            call()
