"""Dense second-quantised reference, independent of the oracle's table code.

Builds H restricted to an (Nup, Ndw) sector from the textbook Jordan-Wigner
definition c_a |n> = (-1)^(sum_{b<a} n_b) |n - e_a>, spin-orbital order
0..L-1 (up) then L..2L-1 (down).  Used only for tiny sectors (dim <= ~5000).
"""
import itertools

import numpy as np


def half_states(L, N):
    return sorted(sum(1 << i for i in c) for c in itertools.combinations(range(L), N))


def sector_states(L, Nup, Ndw):
    up = half_states(L, Nup)
    dn = half_states(L, Ndw)
    # label = i_up + dim_up * i_dn  <=>  word = up | dn << L (spin-orbital bits)
    return [u | (d << L) for d in dn for u in up]


def _annihilate(word, a):
    if not (word >> a) & 1:
        return None, 0
    sign = -1 if bin(word & ((1 << a) - 1)).count("1") % 2 else 1
    return word ^ (1 << a), sign


def _create(word, a):
    if (word >> a) & 1:
        return None, 0
    sign = -1 if bin(word & ((1 << a) - 1)).count("1") % 2 else 1
    return word | (1 << a), sign


def dense_hamiltonian(L, terms, Nup, Ndw):
    """terms: [(type, value, [(r, c, v), ...])] with the reference's element
    conventions (one-body Hermitian-completed from r <= c, the lower triangle
    the conjugate of the upper one; interaction n_r n_c).  Complex values give
    a complex matrix."""
    states = sector_states(L, Nup, Ndw)
    index = {s: i for i, s in enumerate(states)}
    n = len(states)
    cx = any(isinstance(v, complex) and v.imag != 0.0 for (_, _, els) in terms for (_, _, v) in els)
    H = np.zeros((n, n), dtype=np.complex128 if cx else np.float64)
    for (typ, val, els) in terms:
        typ = {"one-body": 0, "interaction": 1}.get(typ, typ)
        for (r, c, v) in els:
            if typ == 1:
                for i, s in enumerate(states):
                    if (s >> r) & 1 and (s >> c) & 1:
                        H[i, i] += val * v
                continue
            pairs = [(r, c, v)] if r == c else ([(r, c, v), (c, r, np.conj(v))] if r < c else [])
            for (a, b, w) in pairs:
                for i, s in enumerate(states):
                    s1, g1 = _annihilate(s, b)
                    if s1 is None:
                        continue
                    s2, g2 = _create(s1, a)
                    if s2 is None or s2 not in index:
                        continue
                    H[index[s2], i] += val * w * g1 * g2
    return H


def dense_c(L, a, Nup, Ndw, dagger):
    """Matrix of c_a (dagger=False) or c^dagger_a from sector (Nup, Ndw)."""
    spin = 1 if a >= L else 0
    d = 1 if dagger else -1
    tNup, tNdw = Nup + (d if spin == 0 else 0), Ndw + (d if spin == 1 else 0)
    src = sector_states(L, Nup, Ndw)
    dst = sector_states(L, tNup, tNdw)
    index = {s: i for i, s in enumerate(dst)}
    M = np.zeros((len(dst), len(src)))
    for j, s in enumerate(src):
        s2, g = (_create(s, a) if dagger else _annihilate(s, a))
        if s2 is None:
            continue
        M[index[s2], j] = g
    return M
