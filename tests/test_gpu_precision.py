"""Reconstruction error of the decomposition (gi_chol_residual) and the
precision form: split-BF16 products (four 8-bit slices per element, exact
FP32 sums per 128 k, FP64 across chunks) for the MO transform
(GI_MO_SPLIT_BF16) and the Step-1 trailing update (GI_CHOL_SPLIT_BF16),
against the FP64 path."""
import numpy as np
import pytest

from gigpu import lib, systems

pytestmark = pytest.mark.gpu
TAU = 1e-6


@pytest.fixture(scope="module")
def medium():
    basis = systems.water_system(3, 6.0, "large", seed=4)
    ctx = lib.Context(0)
    ctx.set_basis(basis)
    yield ctx, basis
    ctx.close()


def residual_host(ctx, basis):
    """(pq|pq) - sum_k R_k(pq)^2 from the host copies of Vdiag and Rkpq."""
    meta = ctx.chol_meta()
    R = ctx.chol_rkpq()                      # (nsub, msd, nvec)
    vd = ctx.chol_vdiag()
    nsh = basis.nshells
    n = [basis.shell_nang(s) for s in range(nsh)]
    vdloc, loc = {}, 0
    for sb in range(nsh):
        for sa in range(sb, nsh):
            vdloc[(sa, sb)] = loc
            loc += n[sa] * n[sb] if sa != sb else n[sa] * (n[sa] + 1) // 2
    out = []
    for s in range(meta["nshell_pairs"]):
        A, B = meta["shell_pairs"][s] - 1
        pos, X = meta["shell_pair_loc"][s][1] - 1, meta["shell_pair_loc"][s][2] - 1
        for q in range(meta["shell_pair_dim"][s]):
            out.append(vd[vdloc[(A, B)] + q] - np.dot(R[X, pos + q], R[X, pos + q]))
    return np.array(out)


def test_residual_matches_host_and_is_bounded(medium):
    ctx, basis = medium
    ctx.chol_build(tau=TAU, flags=lib.CHOL_DECODE_FIXED)
    mx, mn = ctx.chol_residual()
    r = residual_host(ctx, basis)
    assert abs(mx - r.max()) <= 1e-12 and abs(mn - r.min()) <= 1e-12
    # the decomposition of the Step-1 pivots leaves every residual diagonal
    # below tau, and V - R^T R is positive semidefinite
    assert mx <= TAU * (1 + 1e-9)
    assert mn >= -1e-10


def test_reference_decode_leaves_larger_residual(medium):
    """chol2_pq2p_ge_q's traversal (TwoStepCholesky_Step2.f90:520-530)
    factorizes a slightly different pivot set than Step 1 chose; the
    residual then exceeds tau while staying positive semidefinite."""
    ctx, _ = medium
    ctx.chol_build(tau=TAU)
    mx, mn = ctx.chol_residual()
    ctx.chol_build(tau=TAU, flags=lib.CHOL_DECODE_FIXED)
    mx_fixed, _ = ctx.chol_residual()
    assert mx >= mx_fixed
    assert mn >= -1e-10


def test_split_mo_transform_error(medium):
    ctx, basis = medium
    ctx.chol_build(tau=TAU)
    rng = np.random.default_rng(2)
    Cm, _ = np.linalg.qr(rng.normal(size=(basis.nao, basis.nao)))
    nocc, nvirt = 10, 60
    r64 = ctx.mo_transform(Cm, 1, nocc, Cm, nocc + 1, nocc + nvirt)
    r16 = ctx.mo_transform(Cm, 1, nocc, Cm, nocc + 1, nocc + nvirt, flags=lib.MO_SPLIT_BF16)
    # four 8-bit slices per element, exact FP32 sums per 128 k: the error is
    # the slice truncation, ~2^-32 (max|W| sum|C| + ...) per element;
    # measured 1.8e-9 of max |R| (FP32 accumulation gave 8e-8)
    err = np.abs(r16 - r64).max()
    print("split MO transform: max abs err %.3e (max |R| %.3e)" % (err, np.abs(r64).max()))
    assert err <= 1e-8 * np.abs(r64).max()
    assert err > 0.0                              # it is the split path


def test_split_trailing_update_keeps_a_valid_decomposition(medium):
    ctx, basis = medium
    ctx.chol_build(tau=TAU, flags=lib.CHOL_DECODE_FIXED)
    nv64 = ctx.chol_dims()["nvecs"]
    ctx.chol_build(tau=TAU, flags=lib.CHOL_DECODE_FIXED | lib.CHOL_SPLIT_BF16)
    meta = ctx.chol_meta()
    piv = meta["pivot_order"]
    assert len(np.unique(piv)) == len(piv) == meta["nvecs"]
    mx, mn = ctx.chol_residual()
    print("split trailing update: nvecs %d (FP64 %d), residual max %.3e min %.3e" % (meta["nvecs"], nv64, mx, mn))
    # the split update perturbs the residual diagonals by ~1e-9, far below
    # tau: the decomposition keeps (nearly) the FP64 pivot count and stays
    # within tau with a PSD residual
    assert abs(meta["nvecs"] - nv64) <= max(2, 0.002 * nv64)
    assert mn >= -1e-10
    assert mx <= TAU * 1.01
