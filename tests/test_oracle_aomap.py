"""Pins of the AO-ordering oracle by what the reference's own comments and
conventions fix, independently of its numeric tables:

* the Cartesian Molpro tables against the component names the reference
  lists beside them (Auto2eInterface.f90:153-209) looked up in the Auto2e
  Cartesian order (lx = l..0, ly = l-lx..0);
* the spherical Molpro/Orca tables against the m labels of the comments
  (d0 d2- d1+ ..., :29-66) placed in the Auto2e order m = -l..l;
* every map is a permutation, Dalton/PySCF are the identity, the Open Molcas
  map is the sort of an atom's functions by (l, m, shell) (:102-128);
* the Orca signs sit at m = +-3 for f, g, h and m = +-4 for g, h (:517-562).
"""
import numpy as np
import pytest

import aomap as A
from gigpu.basis import cart_components
from gigpu import systems


def _named(names, l):
    comps = cart_components(l)
    out = []
    for nm in names.split():
        c = (nm.count("x"), nm.count("y"), nm.count("z"))
        assert sum(c) == l
        out.append(comps.index(c) + 1)
    return out


def test_cartesian_molpro_tables_match_component_names():
    assert A.CART_MOLPRO[2] == _named("xx yy zz xy xz yz", 2)
    assert A.CART_MOLPRO[3] == _named("xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz", 3)
    assert A.CART_MOLPRO[4] == _named("xxxx yyyy zzzz xxxy xxxz xyyy yyyz xzzz yzzz xxyy xxzz yyzz xxyz xyyz xyzz", 4)


def _mlabels(labels, l):
    out = []
    for lab in labels.split():
        m = int(lab[1:].rstrip("+-") or 0)
        m = -m if lab.endswith("-") else m
        out.append(m + l + 1)
    return out


@pytest.mark.parametrize("ordering,l,labels", [
    (A.MOLPRO, 2, "d0 d2- d1+ d2+ d1-"),
    (A.MOLPRO, 3, "f1+ f1- f0 f3+ f2- f3- f2+"),
    (A.MOLPRO, 4, "g0 g2- g1+ g4+ g1- g2+ g4- g3+ g3-"),
    (A.MOLPRO, 5, "h1+ h1- h2+ h3+ h4- h3- h4+ h5- h0 h5+ h2-"),
    (A.ORCA, 2, "d0 d1+ d1- d2+ d2-"),
    (A.ORCA, 3, "f0 f1+ f1- f2+ f2- f3+ f3-"),
    (A.ORCA, 4, "g0 g1+ g1- g2+ g2- g3+ g3- g4+ g4-"),
])
def test_spherical_tables_match_m_labels(ordering, l, labels):
    assert A.SPHER[ordering][l] == _mlabels(labels, l)


def test_p_functions():
    # Auto2e keeps p Cartesian (x, y, z); Orca stores (z, x, y)
    assert A.SPHER[A.ORCA][1] == [3, 1, 2]
    for o in (A.MOLPRO, A.DALTON, A.PYSCF, A.OPEN_MOLCAS):
        assert A.SPHER[o][1] == [1, 2, 3]


@pytest.fixture(scope="module")
def shells():
    b = systems.water_system(2, 6.0, "large", seed=71, contracted=True)
    return A.basis_shells(b), b.nao


@pytest.mark.parametrize("ordering", [A.MOLPRO, A.ORCA, A.DALTON, A.OPEN_MOLCAS, A.PYSCF])
def test_maps_are_permutations(shells, ordering):
    (ls, atoms), nao = shells
    m, ph = A.ao_map(ls, atoms, True, ordering)
    assert sorted(m.tolist()) == list(range(1, nao + 1))
    assert set(np.unique(ph)) <= {-1.0, 1.0}
    if ordering in (A.DALTON, A.PYSCF):
        assert np.array_equal(m, np.arange(1, nao + 1))


def test_open_molcas_is_a_sort_by_atom_l_m_shell(shells):
    (ls, atoms), nao = shells
    keys = []
    loc = 1
    for s, l in enumerate(ls):
        for mi in range(2 * l + 1):
            keys.append(((atoms[s], l, mi, s), loc + mi))
        loc += 2 * l + 1
    expect = [pos for _, pos in sorted(keys)]
    m, _ = A.ao_map(ls, atoms, True, A.OPEN_MOLCAS)
    assert m.tolist() == expect


def test_orca_phases_at_m_plus_minus_3_and_4(shells):
    (ls, atoms), nao = shells
    _, ph = A.ao_map(ls, atoms, True, A.ORCA)
    loc = 0
    for l in ls:
        for mi in range(2 * l + 1):
            m = mi - l
            want = -1.0 if (abs(m) == 3 and l >= 3) or (abs(m) == 4 and l >= 4) else 1.0
            assert ph[loc + mi] == want
        loc += 2 * l + 1
    assert any(l == 3 for l in ls)


def test_rejected_orderings(shells):
    (ls, atoms), _ = shells
    with pytest.raises(A.OrderingError):
        A.ao_map(ls, atoms, True, A.AUTO2E)
    for o in (A.ORCA, A.DALTON, A.OPEN_MOLCAS, A.PYSCF, A.AUTO2E):
        with pytest.raises(A.OrderingError):
            A.ao_map(ls, atoms, False, o)
    m, ph = A.ao_map(ls, atoms, False, A.MOLPRO)
    nao_cart = sum((l + 1) * (l + 2) // 2 for l in ls)
    assert sorted(m.tolist()) == list(range(1, nao_cart + 1)) and np.all(ph == 1.0)
