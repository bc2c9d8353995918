"""Point-group generators at the model boundary (host side, no GPU work):
the group built from the generators (elements, real characters, symmetric
orbitals) against an independent construction, and the reference's error
cases (invalid or non-commuting generators, sectors outside the group,
operators that break the symmetry)."""
import itertools

import numpy as np
import pytest

from pyqcm_amd import _lib, qcm

# tests/test_diagonalization/S_format.ref:2 (1-based sites, bath phases)
SFORMAT_GENERATORS = [[2, 1, 4, 3, 0, 0, 0, 0, 2, 2, 2, 2], [3, 4, 1, 2, 0, 0, 2, 2, 0, 0, 2, 2]]
_n = [0]


def _name(tag):
    _n[0] += 1
    return "grp_%s_%d" % (tag, _n[0])


def _ring_model(gens):
    name = _name("ring")
    qcm.new_model(name, 4, 0, generators=gens)
    qcm.new_operator(name, "t", "one-body", [(1, 2, -1.0), (2, 3, -1.0), (3, 4, -1.0), (1, 4, -1.0),
                                             (5, 6, -1.0), (6, 7, -1.0), (7, 8, -1.0), (5, 8, -1.0)])
    qcm.new_operator(name, "U", "interaction", [(i, i + 4, 1.0) for i in range(1, 5)])
    return name


def test_sformat_group_sectors():
    name = _name("c2v")
    qcm.new_model(name, 4, 8, generators=SFORMAT_GENERATORS, bath_irrep=True)
    qcm.new_operator(name, "U", "interaction", [(i, i + 12, 1.0) for i in range(1, 5)])
    for r in range(4):
        qcm.new_model_instanceC(name, {"U": 4.0}, "R%d:N12:S0" % r, 500 + r)
    with pytest.raises(_lib.EDError) as e:
        qcm.new_model_instanceC(name, {"U": 4.0}, "R4:N12:S0", 505)
    assert e.value.code == 2


def test_generator_errors():
    with pytest.raises(_lib.EDError):            # not a permutation
        qcm.new_model(_name("bad"), 4, 0, generators=[[1, 1, 3, 4]])
    with pytest.raises(_lib.EDError):            # generators that do not commute
        qcm.new_model(_name("nc"), 3, 0, generators=[[2, 1, 3], [1, 3, 2]])
    with pytest.raises(_lib.EDError) as e:       # order 4: complex characters
        qcm.new_model(_name("c4"), 4, 0, generators=[[2, 3, 4, 1]])
    assert e.value.code == 8
    with pytest.raises(_lib.EDError):            # wrong length
        qcm.new_model(_name("len"), 4, 0, generators=[[2, 1, 4]])


def test_operator_breaking_the_group_is_rejected():
    name = _ring_model([[3, 4, 1, 2], [2, 1, 4, 3]])
    qcm.new_operator(name, "h1", "one-body", [(1, 1, 1.0), (5, 5, 1.0)])   # field on site 1 only
    qcm.new_model_instanceC(name, {"t": 1.0, "U": 4.0}, "R0:N4:S0", 510)
    with pytest.raises(_lib.EDError) as e:
        qcm.new_model_instanceC(name, {"t": 1.0, "U": 4.0, "h1": 0.5}, "R0:N4:S0", 511)
    assert e.value.code == 1


def group_elements(gens, n):
    """elements and real characters of the group of commuting order-2 (or 1)
    site permutations, numbered as products g_0^k0 g_1^k1 ... (k0 fastest)"""
    orders = []
    for g in gens:
        p = list(range(n))
        k = 0
        while True:
            p = [g[i] for i in p]
            k += 1
            if p == list(range(n)):
                break
        orders.append(k)
    elems, digits = [], []
    for ks in itertools.product(*[range(o) for o in reversed(orders)]):
        ks = list(reversed(ks))
        p = list(range(n))
        for g, k in zip(gens, ks):
            for _ in range(k):
                p = [g[i] for i in p]
        elems.append(p)
        digits.append(ks)
    g = len(elems)
    chi = np.array([[np.cos(2 * np.pi * sum(r[j] * k[j] * (g // orders[j]) for j in range(len(orders))) / g)
                     for k in digits] for r in digits])
    return elems, chi


def test_group_characters_are_orthogonal():
    elems, chi = group_elements([[2, 3, 0, 1], [1, 0, 3, 2]], 4)
    assert len(elems) == 4
    assert np.allclose(chi @ chi.T, 4 * np.eye(4))
