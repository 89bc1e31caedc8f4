"""GPU parity on contracted and Cartesian bases.

The other GPU parity tests use uncontracted spherical shells, where every
primitive-pair loop runs once and the Cartesian-to-spherical transform is
always applied.  Here segmented contractions (3- and 2-primitive shells
among single primitives) exercise the primitive-pair accumulation of the ERI
kernels, the contracted Schwarz diagonal and the pair tables, in spherical
and Cartesian AOs; ERIs to 1e-12 absolute, the decomposition exactly in its
pivots and to ERI accuracy in its vectors, the MO transform to 1e-10."""
import numpy as np
import pytest

from gigpu import systems
from gigpu.eri_gen import canonical_classes
from pyoracle import full_vectors
from test_gpu_chol import _compare

pytestmark = pytest.mark.gpu

TOL = 1e-12
BASES = [("spherical", True), ("cartesian", False)]


def _basis(n, size, seed, spher):
    return systems.water_system(n, 6.0, size, seed=seed, contracted=True, spher_ao=spher)


def _nprim(b, s):
    return int(b.nprimitives[b.shell_params_idx[s] - 1])


@pytest.mark.parametrize("name,spher", BASES, ids=[n for n, _ in BASES])
def test_contracted_eri_classes(gpu_ctx, oracle_mod, name, spher):
    """Every canonical class up to (ff|ff) in two caller orientations, with
    contracted shells placed in alternating positions of the quartet."""
    b = _basis(2, "large", 41, spher)
    o = oracle_mod.Oracle(b)
    gpu_ctx.set_basis(b)
    multi = {l: [i for i in range(b.nshells) if b.shell_l(i) == l and _nprim(b, i) > 1] for l in range(4)}
    anyl = {l: [i for i in range(b.nshells) if b.shell_l(i) == l] for l in range(4)}
    assert all(multi[l] for l in range(4))
    rng = np.random.default_rng(3)
    quartets = []
    for k, (la, lb, lc, ld) in enumerate(canonical_classes(3)):
        for perm in ((la, lb, lc, ld), (lc, ld, lb, la)):
            quartets.append([int(rng.choice(multi[l] if (j + k) % 2 == 0 else anyl[l]))
                             for j, l in enumerate(perm)])
    worst = 0.0
    for q, G in zip(quartets, gpu_ctx.eri_quartets(quartets)):
        err = float(np.abs(G - o.eri(*q)).max())
        worst = max(worst, err)
        assert err <= TOL, (q, err)
    print("%s: %d quartets, max |gpu - oracle| = %.2e" % (name, len(quartets), worst))


@pytest.mark.parametrize("md4", [1, 0], ids=["md4", "md_kernel"])
@pytest.mark.parametrize("name,spher", BASES, ids=[n for n, _ in BASES])
def test_contracted_pair_block(gpu_ctx, oracle_mod, name, spher, md4):
    """The product launches (4x4x4 column tiles with ket and bra primitive
    loops, 16x16x4 tiles with stacked kets, lane kernels, transposed
    classes) on contracted shell pairs."""
    b = _basis(2, "large", 43, spher)
    o = oracle_mod.Oracle(b)
    gpu_ctx.set_basis(b)
    rng = np.random.default_rng(9)
    pairs = lambda n: [tuple(sorted(map(int, rng.integers(0, b.nshells, 2)), reverse=True)) for _ in range(n)]
    rp, cp = pairs(31), pairs(29)
    try:
        gpu_ctx.set_option("eri_min_tiles", 1)
        gpu_ctx.set_option("eri_md4", md4)
        V = gpu_ctx.eri_pair_block(rp, cp)
    finally:
        gpu_ctx.set_option("eri_min_tiles", 8192)
        gpu_ctx.set_option("eri_md4", 1)
    nang = b.shell_nang
    r0 = 0
    for (A, B) in rp:
        nr = nang(A) * nang(B)
        c0 = 0
        for (Cs, D) in cp:
            nc = nang(Cs) * nang(D)
            ref = o.eri(A, B, Cs, D).transpose(1, 0, 3, 2).reshape(nr, nc)
            assert np.abs(V[r0:r0 + nr, c0:c0 + nc] - ref).max() <= TOL, (A, B, Cs, D)
            c0 += nc
        r0 += nr


@pytest.mark.parametrize("name,spher,flags", [("spherical", True, 0), ("spherical", True, 1),
                                              ("cartesian", False, 1)],
                         ids=["spherical", "spherical-step1-decode", "cartesian-step1-decode"])
def test_contracted_cholesky(gpu_ctx, oracle_mod, name, spher, flags):
    """Cartesian d shells make the reference's Step-2 decode factorize a
    singular pivot block (d_xx d_yy and d_xy d_xy are proportional), so the
    Cartesian case runs with the Step-1 decode."""
    b = _basis(2, "small", 47, spher)
    _compare(gpu_ctx, oracle_mod.Oracle(b), b, 1e-6, flags=flags)


def test_cartesian_mo_transform(gpu_ctx, oracle_mod):
    b = _basis(2, "small", 53, False)
    o = oracle_mod.Oracle(b)
    gpu_ctx.set_basis(b)
    gpu_ctx.chol_build(tau=1e-6, flags=1)
    ref = o.chol(tau=1e-6, flags=1)
    assert ref.status == 0 and gpu_ctx.chol_dims()["nvecs"] == ref.nvecs
    ref.set_rkpq(gpu_ctx.chol_rkpq())
    rng = np.random.default_rng(2)
    C, _ = np.linalg.qr(rng.normal(size=(b.nao, b.nao)))
    C = np.asfortranarray(C)
    for (a0, a1, b0, b1) in ((1, 10, 11, 40), (3, 12, 1, 5)):
        R = gpu_ctx.mo_transform(C, a0, a1, C, b0, b1)
        Rr = ref.mo_transform(C, a0, a1, C, b0, b1)
        assert np.abs(R - Rr).max() <= 1e-10 * max(1.0, np.abs(Rr).max())
    # the Cartesian decomposition reproduces its factorized columns
    m = gpu_ctx.chol_meta()
    Rg = full_vectors(m["shell_pair_loc"], m["shell_pair_dim"], gpu_ctx.chol_rkpq(), m["norb_pairs"])
    assert np.all(np.isfinite(Rg))


def _deep_basis(spher):
    """Shells contracted from five to seven primitives (more ket primitive
    steps than one Boys pass covers) beside single primitives, s through f."""
    from gigpu.basis import build_basis
    rng = np.random.default_rng(17)

    def shell(l, n, a0):
        ex = [a0 * 2.7 ** k for k in range(n)][::-1]
        return (l, ex, list(rng.uniform(0.2, 0.8, n)))
    shells = {"O": [shell(0, 7, 0.15), shell(0, 1, 0.3), shell(1, 6, 0.2), shell(1, 1, 0.4), shell(2, 5, 0.3),
                    shell(3, 1, 0.6)],
              "H": [shell(0, 5, 0.1), shell(1, 1, 0.5)]}
    coords = np.array([[0.0, 0.0, 0.0], [1.43, 1.1, 0.0], [-1.43, 1.1, 0.2], [0.3, -0.2, 3.1]])
    return build_basis(["O", "H", "H", "O"], coords, shells, spher_ao=spher)


@pytest.mark.parametrize("name,spher", BASES, ids=[n for n, _ in BASES])
def test_deeply_contracted_pair_block(gpu_ctx, oracle_mod, name, spher):
    """Product launches on shells of five to seven primitives: the ket
    primitive steps run in more than one Boys pass."""
    b = _deep_basis(spher)
    o = oracle_mod.Oracle(b)
    gpu_ctx.set_basis(b)
    ns = b.nshells
    allp = [(A, B) for A in range(ns) for B in range(A + 1)]
    try:
        gpu_ctx.set_option("eri_min_tiles", 1)
        V = gpu_ctx.eri_pair_block(allp, allp)
    finally:
        gpu_ctx.set_option("eri_min_tiles", 8192)
    nang = b.shell_nang
    r0 = 0
    worst = 0.0
    for (A, B) in allp:
        nr = nang(A) * nang(B)
        c0 = 0
        for (Cs, D) in allp:
            nc = nang(Cs) * nang(D)
            ref = o.eri(A, B, Cs, D).transpose(1, 0, 3, 2).reshape(nr, nc)
            err = float(np.abs(V[r0:r0 + nr, c0:c0 + nc] - ref).max())
            worst = max(worst, err)
            assert err <= TOL, (A, B, Cs, D, err)
            c0 += nc
        r0 += nr
    print("%s: %d x %d pairs, max |gpu - oracle| = %.2e" % (name, len(allp), len(allp), worst))
