"""Point-group sectors on the device: the state of sector R{r}:N:S is the lowest
state of the full (N, S) sector in representation r (projected Lanczos), its
Green function that of the full target sectors, returned in the site basis or
in the symmetric-orbital basis (blocks).

Checked against an independent dense construction (the group's many-body
operators from the Jordan-Wigner reordering sign, the projected Hamiltonian's
lowest eigenvalue and its Lehmann sum) on a 4-site ring, and against the
reference's C2v fixture (tests/test_diagonalization/S_format.ref) created with
its own generators: E0 and the four representations' Green functions."""
import json
import os

import numpy as np
import pytest

from pyqcm_amd import _lib, qcm, workloads as W
from tests.dense_ref import dense_c, dense_hamiltonian, sector_states
from tests.test_group import SFORMAT_GENERATORS, group_elements

pytestmark = pytest.mark.gpu

_GENS = [[3, 4, 1, 2], [2, 1, 4, 3]]   # 1-based: rotation by two sites, reflection
_n = [0]


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    qcm.set_global_parameterC("GF_method", "L")
    yield


def _ring(U=4.0, t=1.0):
    _n[0] += 1
    name = "gring_%d" % _n[0]
    qcm.new_model(name, 4, 0, generators=_GENS)
    el = [(1, 2, -1.0), (2, 3, -1.0), (3, 4, -1.0), (1, 4, -1.0)]
    qcm.new_operator(name, "t", "one-body", el + [(r + 4, c + 4, v) for (r, c, v) in el])
    qcm.new_operator(name, "U", "interaction", [(i, i + 4, 1.0) for i in range(1, 5)])
    terms = [("one-body", t, [(r - 1, c - 1, v) for (r, c, v) in el + [(r + 4, c + 4, v) for (r, c, v) in el]]),
             ("interaction", U, [(i, i + 4, 1.0) for i in range(4)])]
    return name, {"t": t, "U": U}, terms


def _dense_group_op(perm, L, Nup, Ndw):
    """U(e) on the (Nup, Ndw) sector: c+_a -> c+_{e(a)} on both spins, sign of the
    reordering of the occupied images (up block before down block)"""
    states = sector_states(L, Nup, Ndw)
    index = {s: i for i, s in enumerate(states)}
    U = np.zeros((len(states), len(states)))
    for j, w in enumerate(states):
        occ = [a for a in range(2 * L) if (w >> a) & 1]
        img = [(perm[a % L] + (a // L) * L) for a in occ]
        inv = sum(1 for x in range(len(img)) for y in range(x + 1, len(img)) if img[x] > img[y])
        w2 = sum(1 << b for b in img)
        U[index[w2], j] = -1.0 if inv % 2 else 1.0
    return U


def _dense_irrep_state(terms, L, Nup, Ndw, r):
    elems, chi = group_elements([[v - 1 for v in g] for g in _GENS], L)
    P = sum(chi[r, k] * _dense_group_op(e, L, Nup, Ndw) for k, e in enumerate(elems)) / len(elems)
    w, V = np.linalg.eigh(0.5 * (P + P.T))
    B = V[:, w > 0.5]
    H = dense_hamiltonian(L, terms, Nup, Ndw)
    e, X = np.linalg.eigh(B.T @ H @ B)
    return e[0], B @ X[:, 0]


def test_ring_irrep_ground_states_and_green_functions():
    name, vals, terms = _ring()
    L = 4
    z = W.matsubara(24, 20.0)
    for r in range(4):
        lab = 9300 + r
        qcm.new_model_instanceC(name, vals, "R%d:N4:S0" % r, lab)
        E0, sec = qcm.ground_state_solve(lab)
        assert sec.startswith("R%d:N4:S0" % r)
        E0x, gs = _dense_irrep_state(terms, L, 2, 2, r)
        assert abs(E0 - E0x) < 1e-10 * max(1.0, abs(E0x)), (r, E0, E0x)
        # the Green function of that state through the full target sectors
        G = qcm.Green_function_grid(z, 0, lab)
        Gx = np.zeros_like(G)
        for pm in (1, -1):
            e, V = np.linalg.eigh(dense_hamiltonian(L, terms, 2 + pm, 2))
            A = np.array([V.T @ (dense_c(L, a, 2, 2, pm == 1) @ gs) for a in range(4)])
            den = (z[:, None] - (e - E0x)[None, :]) if pm == 1 else (z[:, None] + (e - E0x)[None, :])
            Gx += np.einsum("ai,bi,zi->zab", A, A, 1.0 / den)
        assert np.abs(G - Gx).max() < 1e-8, (r, np.abs(G - Gx).max())
        # the symmetric-orbital form: S G S^T, blocks of each representation
        Gb = qcm.Green_function_grid(z, 0, lab, blocks=1)
        elems, chi = group_elements([[v - 1 for v in g] for g in _GENS], L)
        rows = []
        for rr in range(4):
            orb = np.zeros(4)
            for k, e in enumerate(elems):
                orb[e[0]] += chi[rr, k]
            if np.linalg.norm(orb) > 1e-9:
                rows.append(orb / np.linalg.norm(orb))
        S = np.array(rows)
        assert np.abs(Gb - np.einsum("ai,zij,bj->zab", S, G, S)).max() < 1e-10


def test_sformat_model_with_its_generators():
    fix = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "s_format_c2v.json")))
    green = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "s_format_c2v_green.json")))
    name = "gc2v_2x2_8b"
    qcm.new_model(name, 4, 8, generators=SFORMAT_GENERATORS, bath_irrep=True)
    for op, d in fix["operators"].items():
        if d["type"] == "one-body":
            qcm.new_operator(name, op, "one-body", [tuple(e) for e in d["elements"]])
    qcm.new_operator(name, "U", "interaction", [(i + 1, i + 13, 1.0) for i in range(4)])
    qcm.new_operator(name, "mu", "one-body", [(i + 1, i + 1, -1.0) for i in range(4)] +
                     [(i + 13, i + 13, -1.0) for i in range(4)])
    P = fix["parameters"]
    vals = {"U": P["U"], "mu": P["mu"]}
    for op, d in fix["operators"].items():
        if d["type"] == "one-body" and P.get(op + "_1", 0.0) != 0.0:
            vals[op] = P[op + "_1"]
    qcm.set_global_parameterC("Qmatrix_wtol", 1e-12)
    try:
        qcm.new_model_instanceC(name, vals, "R0:N12:S0", 9310)
        E0, sec = qcm.ground_state_solve(9310)
        assert sec.startswith("R0:N12:S0")
        assert abs(E0 - fix["E0"]) < 6e-5
        Z = np.array(green["z"]) @ np.array([1.0, 1.0j])
        Gb = qcm.Green_function_grid(Z, 0, 9310, blocks=1)
    finally:
        qcm.set_global_parameterC("Qmatrix_wtol", 1e-5)
    # each representation holds one symmetric d orbital: the diagonal of the
    # block form, against the oracle's converged Green function of that
    # representation on the imaginary axis
    for i in range(4):
        go = np.array(green["G"][i]) @ [1.0, 1.0j]
        d = np.abs(Gb[:, i, i] - go)
        assert d[Z.real == 0.0].max() < 1e-8, (i, d[Z.real == 0.0].max())
    off = Gb - np.einsum("zii->zi", Gb)[:, :, None] * np.eye(4)[None]
    assert np.abs(off).max() < 1e-12
