"""The references of the complex Hilbert-space tests, pinned on the CPU:
  * the fixture read from tests/test_diagonalization/S_format_complex.ref
    (8-site open chain, element +i on each bond, U = 1e-9): its poles are the
    one-body levels +-2 cos(k pi / 9) (shifted by U at the 1e-9 level), its
    Q-matrix reproduces the free-fermion G(z) = (z - h)^-1 with the
    reference's convention G(r, c) = sum_i Q(c, i) conj(Q(r, i)) / (z - e_i),
    and its E0 is twice the sum of the negative levels;
  * the dense complex Jordan-Wigner Hamiltonian (tests/dense_ref.py): Hermitian,
    at U = 0 its ground energy is twice the sum of the negative one-body
    levels, and a complex hopping on an open chain has the spectrum of the real
    chain (gauge transformation), with and without U."""
import json
import os

import numpy as np

from tests.dense_ref import dense_hamiltonian

_FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "s_format_complex.json")))


def _fixture_q():
    q = _FIX["qmatrix"]
    e = np.array(q["poles"])
    Q = np.array([[complex(*x) for x in row] for row in q["residues"]]).T  # Q[a, i]
    return e, Q


def _chain_h(L, amp):
    h = np.zeros((L, L), complex)
    for j in range(L - 1):
        h[j, j + 1] = amp[j]
        h[j + 1, j] = np.conj(amp[j])
    return h


def test_fixture_is_the_free_chain():
    e, Q = _fixture_q()
    L = 8
    h = _chain_h(L, [1j] * (L - 1))
    eps = np.linalg.eigvalsh(h)
    k = np.arange(1, L + 1)
    assert np.abs(np.sort(eps) - np.sort(2 * np.cos(k * np.pi / (L + 1)))).max() < 1e-12
    assert np.abs(np.sort(e) - np.sort(eps)).max() < 1e-8
    assert abs(_FIX["E0"] - 2 * eps[eps < 0].sum()) < 5e-6
    for z in (0.3 + 0.7j, -1.1 + 0.05j, 2j):
        G = np.einsum("ci,ri,i->rc", Q, Q.conj(), 1.0 / (z - e))
        assert np.abs(G - np.linalg.inv(z * np.eye(L) - h)).max() < 1e-8
        # the transposed convention would not match: the residues are complex
        assert np.abs(G.T - np.linalg.inv(z * np.eye(L) - h)).max() > 1e-2


def _terms(L, amp, U, mu):
    hop = [(j, j + 1, amp[j]) for j in range(L - 1)] + [(j + L, j + 1 + L, amp[j]) for j in range(L - 1)]
    return [("one-body", 1.0, hop),
            ("interaction", U, [(j, j + L, 1.0) for j in range(L)]),
            ("one-body", -mu, [(j, j, 1.0) for j in range(2 * L)])]


def test_dense_complex_hamiltonian():
    rng = np.random.default_rng(11)
    L = 5
    phases = rng.uniform(0, 2 * np.pi, L - 1)
    amp_c = [complex(-np.cos(p), -np.sin(p)) for p in phases]
    H0 = dense_hamiltonian(L, _terms(L, amp_c, 0.0, 0.0), 3, 2)
    assert H0.dtype == np.complex128
    assert np.abs(H0 - H0.conj().T).max() == 0.0
    eps = np.linalg.eigvalsh(_chain_h(L, amp_c))
    assert abs(np.linalg.eigvalsh(H0)[0] - (eps[:3].sum() + eps[:2].sum())) < 1e-12
    for U in (0.0, 4.0):
        Hc = dense_hamiltonian(L, _terms(L, amp_c, U, 1.3), 3, 2)
        Hr = dense_hamiltonian(L, _terms(L, [-1.0] * (L - 1), U, 1.3), 3, 2)
        assert Hr.dtype == np.float64
        assert np.abs(np.linalg.eigvalsh(Hc) - np.linalg.eigvalsh(Hr)).max() < 1e-11
