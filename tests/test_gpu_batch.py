"""Batched form: independent decompositions run concurrently (one context
and stream each) give exactly what the same decompositions give one at a
time."""
import numpy as np
import pytest

from gigpu import lib, systems

pytestmark = pytest.mark.gpu


def test_batch_equals_sequential():
    mols = systems.batch_systems(nmol=6, seed=5, nmin=1, nmax=4)
    seq = []
    ref_ctx = lib.Context(0)
    try:
        for b in mols:
            ref_ctx.set_basis(b)
            ref_ctx.chol_build(tau=1e-6)
            seq.append((ref_ctx.chol_meta(), ref_ctx.chol_rkpq()))
    finally:
        ref_ctx.close()
    ctxs = [lib.Context(0) for _ in mols]
    try:
        for c, b in zip(ctxs, mols):
            c.set_basis(b)
        status = lib.chol_build_batch(ctxs, tau=1e-6, nworkers=3)
        assert list(status) == [0] * len(mols)
        for c, (m_ref, r_ref) in zip(ctxs, seq):
            m = c.chol_meta()
            assert m["nvecs"] == m_ref["nvecs"]
            assert np.array_equal(m["pivot_order"], m_ref["pivot_order"])
            r = c.chol_rkpq()
            assert r.shape == r_ref.shape
            assert np.abs(r - r_ref).max() <= 1e-12 * max(1.0, np.abs(r_ref).max())
    finally:
        for c in ctxs:
            c.close()


def test_batch_of_mixed_bases_equals_sequential():
    """Uneven entries: uncontracted, contracted and Cartesian bases, large and
    small, in one batch (the per-context work arrays and class mixes differ
    from entry to entry)."""
    mols = [systems.water_system(2, 6.0, "small", seed=81),
            systems.water_system(1, 6.0, "large", seed=82, contracted=True),
            systems.water_system(2, 6.0, "small", seed=83, contracted=True, spher_ao=False),
            systems.water_system(3, 6.0, "small", seed=84, contracted=True)]
    flags = lib.CHOL_DECODE_FIXED   # Cartesian d shells need the Step-1 decode
    seq = []
    ref_ctx = lib.Context(0)
    try:
        for b in mols:
            ref_ctx.set_basis(b)
            ref_ctx.chol_build(tau=1e-6, flags=flags)
            seq.append((ref_ctx.chol_meta(), ref_ctx.chol_rkpq()))
    finally:
        ref_ctx.close()
    ctxs = [lib.Context(0) for _ in mols]
    try:
        for c, b in zip(ctxs, mols):
            c.set_basis(b)
        status = lib.chol_build_batch(ctxs, tau=1e-6, flags=flags, nworkers=4)
        assert list(status) == [0] * len(mols)
        for c, (m_ref, r_ref) in zip(ctxs, seq):
            m = c.chol_meta()
            assert m["nvecs"] == m_ref["nvecs"]
            assert np.array_equal(m["pivot_order"], m_ref["pivot_order"])
            r = c.chol_rkpq()
            assert np.abs(r - r_ref).max() <= 1e-12 * max(1.0, np.abs(r_ref).max())
    finally:
        for c in ctxs:
            c.close()


def test_batch_reports_failing_entry():
    mols = systems.batch_systems(nmol=2, seed=6, nmin=1, nmax=2)
    ctxs = [lib.Context(0) for _ in mols]
    try:
        ctxs[0].set_basis(mols[0])          # the second context has no basis
        with pytest.raises(lib.GigpuError, match="batch entry 1"):
            lib.chol_build_batch(ctxs, tau=1e-6, nworkers=2)
    finally:
        for c in ctxs:
            c.close()


def _fingerprint(c):
    m = c.chol_meta()
    x = c.chol_inv_l()
    return m["nvecs"], m["pivot_order"].copy(), np.diag(x).copy(), float(np.abs(x).sum())


def test_full_batch_repeated_under_load():
    """32 systems on 8 workers, twice: every entry succeeds and matches its
    one-at-a-time build (pivot order exactly, L^-1 to rounding).  The
    factorization of the pivot block runs while other entries' ERI, GEMM and
    pivot-search kernels share the GPU."""
    mols = systems.batch_systems(nmol=32, seed=3, nmin=1, nmax=8)
    ctxs = [lib.Context(0) for _ in mols]
    try:
        for c, b in zip(ctxs, mols):
            c.set_basis(b)
        ref = []
        for c in ctxs:
            c.chol_build(tau=1e-6)
            ref.append(_fingerprint(c))
        for rep in range(2):
            status = lib.chol_build_batch(ctxs, tau=1e-6, nworkers=8)
            assert list(status) == [0] * len(mols)
            for i, (c, (nv, piv, dg, s)) in enumerate(zip(ctxs, ref)):
                nv2, piv2, dg2, s2 = _fingerprint(c)
                assert nv2 == nv, (rep, i)
                assert np.array_equal(piv2, piv), (rep, i)
                assert np.abs(dg2 - dg).max() <= 1e-10 * np.abs(dg).max(), (rep, i)
                assert abs(s2 - s) <= 1e-10 * s, (rep, i)
    finally:
        for c in ctxs:
            c.close()


def test_autotuner_picks_from_candidates_and_keeps_results():
    from gigpu import autotune
    mols = systems.batch_systems(nmol=6, seed=8, nmin=1, nmax=3)
    ctxs = [lib.Context(0) for _ in mols]
    try:
        for c, b in zip(ctxs, mols):
            c.set_basis(b)
        lib.chol_build_batch(ctxs, tau=1e-6, nworkers=2)
        ref = [c.chol_dims()["nvecs"] for c in ctxs]
        cand = {"eri_max_kt": [32, 8], "eri_lane_emax": [4, 2]}
        best, t, trials = autotune.tune(ctxs, candidates=cand, nworkers=[2, 4], reps=1)
        assert best["eri_max_kt"] in (32, 8) and best["eri_lane_emax"] in (4, 2)
        assert best["nworkers"] in (2, 4)
        assert t == min(tt for _, tt in trials)
        assert len(trials) == 4
        lib.chol_build_batch(ctxs, tau=1e-6, nworkers=best["nworkers"])
        assert [c.chol_dims()["nvecs"] for c in ctxs] == ref
    finally:
        for c in ctxs:
            c.close()
