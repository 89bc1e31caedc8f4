"""GPU parity of the ERI class kernels against the oracle: every canonical
class up to (ff|ff), every caller orientation, to 1e-12 absolute."""
import itertools

import numpy as np
import pytest

from gigpu import systems
from gigpu.eri_gen import canonical_classes

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def setup(gpu_ctx, oracle_mod):
    b = systems.water_system(2, 6.0, "large", seed=21)
    gpu_ctx.set_basis(b)
    return b, oracle_mod.Oracle(b), gpu_ctx


def _pick(b, rng, ls):
    by_l = {l: [i for i in range(b.nshells) if b.shell_l(i) == l] for l in range(4)}
    return [int(rng.choice(by_l[l])) for l in ls]


def test_all_classes_all_orientations(setup):
    b, o, ctx = setup
    rng = np.random.default_rng(0)
    quartets = []
    for (la, lb, lc, ld) in canonical_classes(3):
        # the canonical order and every swap the frontends undo
        for perm in {(la, lb, lc, ld), (lb, la, lc, ld), (la, lb, ld, lc), (lc, ld, la, lb), (ld, lc, lb, la)}:
            quartets.append(_pick(b, rng, perm))
    blocks = ctx.eri_quartets(quartets)
    worst = 0.0
    for q, G in zip(quartets, blocks):
        R = o.eri(*q)
        worst = max(worst, float(np.abs(G - R).max()))
        assert np.abs(G - R).max() <= TOL, (q, float(np.abs(G - R).max()))
    print("max |gpu - oracle| =", worst)


def test_same_shell_and_one_centre(setup):
    b, o, ctx = setup
    quartets = []
    for s in range(0, b.nshells, 3):
        quartets.append([s, s, s, s])
        quartets.append([s, (s + 1) % b.nshells, s, (s + 1) % b.nshells])
    for q, G in zip(quartets, ctx.eri_quartets(quartets)):
        assert np.abs(G - o.eri(*q)).max() <= TOL


def test_range_separated_kappa(setup):
    b, o, ctx = setup
    rng = np.random.default_rng(4)
    kappa = 1.0 / 0.5 ** 2
    quartets = [list(rng.integers(0, b.nshells, 4)) for _ in range(60)]
    for q, G in zip(quartets, ctx.eri_quartets(quartets, kappa=kappa)):
        assert np.abs(G - o.eri(*q, kappa=kappa)).max() <= TOL


def test_many_random_quartets(setup):
    b, o, ctx = setup
    rng = np.random.default_rng(8)
    quartets = [list(rng.integers(0, b.nshells, 4)) for _ in range(1500)]
    for q, G in zip(quartets, ctx.eri_quartets(quartets)):
        assert np.abs(G - o.eri(*q)).max() <= TOL


def _pairs(b, rng, lpairs, n):
    by_l = {l: [i for i in range(b.nshells) if b.shell_l(i) == l] for l in range(4)}
    out = []
    for k in range(n):
        l0, l1 = lpairs[k % len(lpairs)]
        out.append((int(rng.choice(by_l[l0])), int(rng.choice(by_l[l1]))))
    return out


ALL_L = [(a, c) for a in range(4) for c in range(4)]
THIN = [(0, 0), (1, 0), (0, 1), (1, 1)]
THICK = [(3, 3), (3, 2), (2, 3), (2, 2), (3, 1)]


@pytest.mark.parametrize("rows,cols", [(ALL_L, ALL_L), (THICK, THIN), (THIN, THICK)],
                         ids=["mixed", "thin-kets", "thin-bras"])
@pytest.mark.parametrize("lane,trans,kt,md4,md4_tiles", [(1, 1, 32, 1, 2048), (1, 1, 32, 1, 0), (1, 1, 32, 1, 6),
                                                         (0, 1, 32, 1, 2048), (0, 0, 32, 1, 2048),
                                                         (0, 1, 32, 0, 2048), (0, 0, 32, 0, 2048),
                                                         (0, 1, 1, 0, 2048)],
                         ids=["default", "md4-four-rows", "md4-two-rows", "md4-only", "md4-no-transpose",
                              "md_kernel", "md_kernel-no-transpose", "md_kernel-one-ket-tiles"])
def test_pair_block_launch_shapes(setup, rows, cols, lane, trans, kt, md4, md4_tiles):
    """The product launches with every kernel choice: the 4x4x4 kernel
    (column tiles of four slots, per-class orientation, row blocks of one,
    two or four bra pairs), the 16x16x4 kernel (ket stacking across pairs of
    one tile, transposed orientation, single-ket tiles), lane kernels for the
    small classes."""
    b, o, ctx = setup
    rng = np.random.default_rng(7)
    rp, cp = _pairs(b, rng, rows, 37), _pairs(b, rng, cols, 41)
    try:
        ctx.set_option("eri_min_tiles", 1)
        ctx.set_option("eri_max_kt", kt)
        ctx.set_option("eri_lane", lane)
        ctx.set_option("eri_trans", trans)
        ctx.set_option("eri_md4", md4)
        ctx.set_option("eri_md4_tiles", md4_tiles)
        V = ctx.eri_pair_block(rp, cp)
    finally:
        ctx.set_option("eri_min_tiles", 8192)
        ctx.set_option("eri_max_kt", 32)
        ctx.set_option("eri_lane", 1)
        ctx.set_option("eri_trans", 1)
        ctx.set_option("eri_md4", 1)
        ctx.set_option("eri_md4_tiles", 2048)
    nang = b.shell_nang
    r0 = 0
    worst = 0.0
    for (A, B) in rp:
        nr = nang(A) * nang(B)
        c0 = 0
        for (Cs, D) in cp:
            nc = nang(Cs) * nang(D)
            # row a + na b, column c + nc d
            ref = o.eri(A, B, Cs, D).transpose(1, 0, 3, 2).reshape(nr, nc)
            worst = max(worst, float(np.abs(V[r0:r0 + nr, c0:c0 + nc] - ref).max()))
            c0 += nc
        r0 += nr
    assert worst <= TOL, worst
