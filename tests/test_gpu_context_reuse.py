"""A context keeps its device work arrays across builds (grow-only: the
Step-1 panels, the Vabrs staging, the factorization scratch and the
split-product slices).  Builds of different systems in any order on one
context must give bit-identical results to fresh contexts: stale contents of
a reused array, or an array sized for an earlier, smaller system, would show
here."""
import numpy as np
import pytest

from gigpu import lib, systems

pytestmark = pytest.mark.gpu


def _build(ctx, basis, flags):
    ctx.set_basis(basis)
    ctx.chol_build(tau=1e-6, flags=flags)
    m = ctx.chol_meta()
    return m["nvecs"], m["pivot_order"].copy(), ctx.chol_rkpq().copy(), ctx.chol_inv_l().copy()


def _fresh(basis, flags):
    c = lib.Context(0)
    try:
        return _build(c, basis, flags)
    finally:
        c.close()


@pytest.mark.parametrize("flags", [0, lib.CHOL_DECODE_FIXED | lib.CHOL_SPLIT_BF16], ids=["fp64", "split"])
def test_builds_on_one_context_match_fresh_contexts(flags):
    small = systems.water_system(1, 6.0, "small", seed=61)
    large = systems.water_system(3, 6.0, "small", seed=62)
    ref_small, ref_large = _fresh(small, flags), _fresh(large, flags)
    ctx = lib.Context(0)
    try:
        for basis, ref in ((small, ref_small), (large, ref_large), (small, ref_small), (large, ref_large)):
            got = _build(ctx, basis, flags)
            assert got[0] == ref[0]
            assert np.array_equal(got[1], ref[1])
            assert np.array_equal(got[2], ref[2])
            assert np.array_equal(got[3], ref[3])
    finally:
        ctx.close()


def test_split_mo_transform_after_a_larger_one():
    """The split-product slices sized by a large MO transform are reused by
    a smaller one and by the trailing update."""
    b = systems.water_system(2, 6.0, "small", seed=63)
    rng = np.random.default_rng(3)
    C, _ = np.linalg.qr(rng.normal(size=(b.nao, b.nao)))
    C = np.asfortranarray(C)
    fresh = lib.Context(0)
    try:
        fresh.set_basis(b)
        fresh.chol_build(tau=1e-6)
        ref = fresh.mo_transform(C, 1, 5, C, 6, 20, flags=lib.MO_SPLIT_BF16)
    finally:
        fresh.close()
    ctx = lib.Context(0)
    try:
        ctx.set_basis(b)
        ctx.chol_build(tau=1e-6)
        ctx.mo_transform(C, 1, 30, C, 31, b.nao, flags=lib.MO_SPLIT_BF16)
        got = ctx.mo_transform(C, 1, 5, C, 6, 20, flags=lib.MO_SPLIT_BF16)
        assert np.array_equal(got, ref)
    finally:
        ctx.close()


def test_build_without_vectors_drops_earlier_vectors():
    """Rkpq keeps its device array across builds; a build that stops after
    Step 2 (GI_CHOL_SKIP_VECTORS) must not leave the earlier vectors
    readable as if they belonged to it."""
    b = systems.water_system(1, 6.0, "small", seed=64)
    ctx = lib.Context(0)
    try:
        ctx.set_basis(b)
        ctx.chol_build(tau=1e-6)
        assert ctx.chol_rkpq().size > 0
        ctx.chol_build(tau=1e-6, flags=lib.CHOL_SKIP_VECTORS)
        with pytest.raises(lib.GigpuError, match="not available"):
            ctx.chol_rkpq()
        ctx.chol_build(tau=1e-6)
        assert ctx.chol_rkpq().size > 0
    finally:
        ctx.close()
