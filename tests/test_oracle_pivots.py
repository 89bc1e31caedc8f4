"""Pins of the oracle's Step-1 pivot search (oracle/chol.c) against the dense
restatement in dense_pivots.py, which is written from the reference's
Fortran (TwoStepCholesky_Step1.f90:175-442,533-730,859-1103) and shares no
code with the oracle: the pivot sequence, the BASE shell pairs and the vector
count must be identical.  The cases cover an uncontracted basis, contracted
shells, a geometry where the Schwarz prescreening (DefineBase) discards shell
pairs and keeps the last discarded one, and same-shell p/d pairs that are
only partly pivoted.  A planted change of any heuristic (the sigma = 0.01
spread, the batch limit, the candidate order, the BASE cut) must break the
agreement, which test_planted_heuristic_changes_are_detected checks on the
restatement itself."""
import numpy as np
import pytest

import dense_pivots
from gigpu import systems
from gigpu.basis import build_basis
from grad_helpers import tiny_basis


def far_apart_basis():
    """Two water-like fragments 14 bohr apart with tight functions: shell
    pairs across the gap have (ab|ab) far below min(tau, 1e-8)^2 / MaxDrs."""
    coords = np.array([[0.0, 0.0, 0.0], [1.4, 1.1, 0.0], [-1.4, 1.1, 0.0], [14.0, 0.0, 0.0]])
    O = [(0, [5.0], [1.0]), (0, [1.2], [1.0]), (1, [2.0], [1.0]), (2, [1.5], [1.0])]
    H = [(0, [3.0], [1.0]), (1, [1.6], [1.0])]
    Ne = [(0, [9.0], [1.0]), (1, [4.0], [1.0]), (2, [2.5], [1.0])]
    return build_basis(["O", "H", "H", "Ne"], coords, {"O": O, "H": H, "Ne": Ne})


CASES = {
    "uncontracted": lambda: systems.water_system(1, 6.0, "small", seed=7),
    "contracted": lambda: tiny_basis(lmax=3, contracted=True),
    "contracted-water": lambda: systems.water_system(1, 6.0, "small", seed=11, contracted=True),
    "prescreened": far_apart_basis,
    # 3240 orbital pairs: more candidates than the 1000-column batch, so the
    # batch limit and the sigma spread shape the sequence
    "two-waters": lambda: systems.water_system(2, 6.0, "small", seed=5),
}


def _dense_all(o, b):
    return o.pair_matrix(dense_pivots.vdiag_shell_pairs(b.nshells))


@pytest.fixture(scope="module", params=sorted(CASES))
def case(oracle_mod, request):
    b = CASES[request.param]()
    o = oracle_mod.Oracle(b)
    return request.param, b, o, _dense_all(o, b)


@pytest.mark.parametrize("tau", [1e-6, 1e-8])
def test_pivot_sequence_equals_dense_restatement(case, tau):
    name, b, o, V = case
    piv, sp, nvec, dmax, conv = dense_pivots.step1(V, b, tau, quantized=True)
    r = o.chol(tau=tau, flags=1 | 2)
    assert r.status == 0, r.message
    assert conv and dmax <= tau
    assert np.array_equal(r.shell_pairs, sp), name
    assert r.nvecs == nvec
    assert np.array_equal(r.pivot_order, piv), name
    if name == "prescreened":
        nsh = b.nshells
        assert len(sp) < nsh * (nsh + 1) // 2     # DefineBase dropped pairs (one more than it kept)


def asymmetric_basis(contracted):
    """O, H, H in a general position with a different basis on every atom, so
    that no two orbital pairs have equal diagonals in exact arithmetic."""
    coords = np.array([[0.0, 0.1, -0.2], [1.4, 1.1, 0.3], [-1.5, 0.9, 0.4]])
    O = [(0, [0.45], [1.0]), (1, [0.8], [1.0]), (2, [1.1], [1.0]), (0, [2.3], [1.0]), (1, [2.9], [1.0])]
    if contracted:
        O += [(0, [6.0, 1.3, 0.35], [0.15, 0.55, 0.45]), (1, [3.0, 0.6], [0.4, 0.7])]
    H1 = [(0, [0.31], [1.0]), (0, [1.5], [1.0]), (1, [0.75], [1.0])]
    H2 = [(0, [0.37], [1.0]), (0, [1.9], [1.0]), (1, [0.66], [1.0])]
    return build_basis(["O", "H", "He"], coords, {"O": O, "H": H1, "He": H2})


@pytest.mark.parametrize("tau", [1e-6, 1e-8])
def test_raw_keys_equal_dense_restatement(oracle_mod, tau):
    """OR_CHOL_RAW_TIES: raw comparisons, the reference's own rule, on a
    basis with a different basis per atom.  Exact ties still exist between
    components of one-centre same-shell pairs ((px px|px px) = (py py|py py));
    two implementations order exactly tied keys by the last bits of their
    different summation orders, which for this uncontracted basis happens not
    to change the sequence (with the contracted O shells added it swaps two
    tied pivots, which is why the library and the oracle quantize keys)."""
    b = asymmetric_basis(False)
    o = oracle_mod.Oracle(b)
    V = _dense_all(o, b)
    piv, sp, nvec, _, _ = dense_pivots.step1(V, b, tau, quantized=False)
    r = o.chol(tau=tau, flags=1 | 2 | 4)
    assert r.status == 0
    assert np.array_equal(r.shell_pairs, sp)
    assert r.nvecs == nvec
    assert np.array_equal(r.pivot_order, piv)


def test_same_shell_pairs_partly_pivoted(oracle_mod):
    """Some same-shell pair of l >= 1 has pivots and non-pivots, so the
    lower-triangle storage order of same-shell pairs is exercised."""
    b = CASES["uncontracted"]()
    o = oracle_mod.Oracle(b)
    r = o.chol(tau=1e-6, flags=1 | 2)
    partial = 0
    for (A, B), (loc, _, _), dim in zip(r.shell_pairs, r.shell_pair_loc, r.shell_pair_dim):
        if A == B and b.shell_l(A - 1) >= 1:
            s = r.pivots[loc - 1:loc - 1 + dim].sum()
            partial += 0 < s < dim
    assert partial > 0


def test_prescreening_keeps_the_last_discarded_pair(oracle_mod):
    """DefineBase keeps shell pair k0, the last one at or below the
    prescreening threshold (TwoStepCholesky_Step1.f90:391-419)."""
    b = far_apart_basis()
    o = oracle_mod.Oracle(b)
    V = _dense_all(o, b)
    r = o.chol(tau=1e-6, flags=1 | 2)
    d = np.diag(V)
    pairs = dense_pivots.vdiag_shell_pairs(b.nshells)
    loc = np.cumsum([0] + [len(x) for x in _pair_rows(b, pairs)])
    vmax = {p: d[loc[i]:loc[i + 1]].max() for i, p in enumerate(pairs)}
    thresh = min(1e-6, 1e-8) ** 2 / max(vmax.values())
    kept = [tuple(x) for x in r.shell_pairs]
    below = [p for p in kept if vmax[p] <= thresh]
    assert len(below) == 1                       # exactly the pair at k0
    assert vmax[below[0]] == max(v for p, v in vmax.items() if v <= thresh)
    assert all(vmax[p] > thresh for p in kept if p != below[0])


def _pair_rows(b, pairs):
    out = []
    for A, B in pairs:
        na, nb = b.shell_nang(A - 1), b.shell_nang(B - 1)
        out.append([0] * (na * nb if A != B else na * (na + 1) // 2))
    return out


@pytest.mark.parametrize("planted,where", [(("SIGMA", 0.001), "two-waters"), (("SIGMA", 0.05), "two-waters"),
                                           (("MIN_BATCH", 400), "two-waters"),
                                           (("KEEP_K0", False), "prescreened")])
def test_planted_heuristic_changes_are_detected(case, monkeypatch, planted, where):
    """The comparison above would catch a wrong spread factor, batch limit or
    BASE cut: each planted in the restatement changes its result.  (With at
    most 1000 orbital pairs the batch holds every candidate and the greedy
    pivot order does not depend on sigma or the batch limit, hence the
    two-water case for those.)"""
    name, b, o, V = case
    if name != where:
        pytest.skip("planted change checked on the %s case" % where)
    attr, value = planted
    piv, sp, *_ = dense_pivots.step1(V, b, 1e-6)
    monkeypatch.setattr(dense_pivots, attr, value)
    piv_b, sp_b, *_ = dense_pivots.step1(V, b, 1e-6)
    assert not (np.array_equal(piv_b, piv) and np.array_equal(sp_b, sp))
