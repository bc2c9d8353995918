"""The numpy lattice-sum oracle (oracle/lattice_ref.py) against closed forms of
the non-interacting limit, where the cluster Green function is (z - h_c)^{-1}
and the lattice is known exactly:
  * the CPT Green function averaged over the Brillouin zone is the local
    lattice Green function, a Bloch sum over eps(q) = -2t cos q - mu;
  * log|det(V(k) G - 1)| = sum_n log|z - eps_n(k)| - log|det(z - h_c)|;
  * the grid Potthoff functional is the lattice grand potential per site,
    2 / (nk L) sum min(eps, 0), whatever Weiss field the cluster carries.
The Bloch energies come from the cosine formula, not from V(k), so a wrong
phase sign, bond, element orientation or spin factor fails these tests."""
import numpy as np
import pytest

from oracle import lattice_ref as LR
from pyqcm_amd import lattice as LT


def _chain_hc(L, t, mu, weiss=None):
    h = np.zeros((L, L))
    for i in range(L - 1):
        h[i, i + 1] = h[i + 1, i] = -t
    h -= mu * np.eye(L)
    if weiss is not None:
        h += np.diag(weiss)
    return h


def _bloch(L, t, mu, ks):
    """eps_m(k) = -2t cos(2 pi (k + m) / L) - mu: the chain's bands folded onto the L-site superlattice."""
    return np.array([[-2 * t * np.cos(2 * np.pi * (k + m) / L) - mu for m in range(L)] for k in ks])


@pytest.mark.parametrize("L,nk", [(4, 16), (3, 7)])
def test_cpt_average_u0_is_the_local_lattice_green_function(L, nk):
    t, mu = 1.0, 0.3
    hc = _chain_hc(L, t, mu)
    z = np.array([-1.3 + 0.1j, 0.2 + 0.05j, 2.7 + 0.3j, 0.9j, 3.1j])
    G0 = np.array([np.linalg.inv(zz * np.eye(L) - hc) for zz in z])
    sl = LT.chain_superlattice(L, t)
    P = LR.cpt_projected_gf(G0, sl, (1, (nk,), 0.5))
    ks = (np.arange(nk) + 0.5) / nk
    eps = _bloch(L, t, mu, ks)
    for iz, zz in enumerate(z):
        exact = np.mean(1.0 / (zz - eps))
        assert abs(np.trace(P[iz]) / L - exact) < 1e-12
        # the full matrix equals the k-average of (z - h(k))^{-1}
        direct = np.mean([np.linalg.inv(zz * np.eye(L) - hc - LR.superlattice_V(sl, (k,))) for k in ks], axis=0)
        assert np.abs(P[iz] - direct).max() < 1e-12


def test_sef_integrand_u0_identity():
    L, t, mu, nk = 4, 1.0, -0.4, 12
    hc = _chain_hc(L, t, mu)
    w = np.array([0.05, 0.7, 2.0, 11.0])
    G0 = np.array([np.linalg.inv(1j * x * np.eye(L) - hc) for x in w])
    sl = LT.chain_superlattice(L, t)
    got = LR.sef_logdet(G0, sl, (1, (nk,), 0.0))
    ks = np.arange(nk) / nk
    eps = _bloch(L, t, mu, ks)
    ec = np.linalg.eigvalsh(hc)
    for n, x in enumerate(w):
        exact = np.mean(np.sum(np.log(np.abs(1j * x - eps)), axis=1)) - np.sum(np.log(np.abs(1j * x - ec)))
        assert abs(got[n] - exact) < 1e-12


@pytest.mark.parametrize("weiss_kind", ["none", "staggered", "neel"])
def test_grid_potthoff_functional_u0_is_the_lattice_grand_potential(weiss_kind):
    L, t, mu, nk = 4, 1.0, 0.35, 32
    w, p = LT.legendre_frequency_grid(1.0, 20.0, 60, 60, 60)
    stag = np.array([(-1.0) ** i for i in range(L)])
    # a uniform shift too, so that Tr V0 != 0 and the diagonal term is tested
    W_up = {"none": np.zeros(L), "staggered": 0.3 * stag + 0.2, "neel": 0.3 * stag + 0.2}[weiss_kind]
    W_dn = -0.3 * stag + 0.2 if weiss_kind == "neel" else W_up
    mixing = weiss_kind == "neel"
    I, E0 = 0.0, 0.0
    for W, spin_factor in ((W_up, 1.0 if mixing else 2.0), (W_dn, 1.0 if mixing else 0.0)):
        if spin_factor == 0.0:
            continue
        hc = _chain_hc(L, t, mu, W)
        G0 = np.array([np.linalg.inv(1j * x * np.eye(L) - hc) for x in w])
        sl = LT.chain_superlattice(L, t, weiss=W)
        I += spin_factor * np.dot(p, LR.sef_logdet(G0, sl, (1, (nk,), 0.0)))
        ec = np.linalg.eigvalsh(hc)
        E0 += spin_factor * np.sum(ec[ec < 0])
    V0_up = LT.chain_superlattice(L, t, weiss=W_up)["V0"]
    V0_dn = LT.chain_superlattice(L, t, weiss=W_dn)["V0"]
    omega = LR.potthoff_functional(I, E0, V0_up, V0_dn, L, mixing)
    eps = _bloch(L, t, mu, np.arange(nk) / nk)
    exact = 2.0 * np.sum(np.minimum(eps, 0.0)) / (nk * L)
    assert abs(omega - exact) < 1e-9, (omega, exact)


def test_frequency_grids_integrate_the_lorentzian():
    """Both grids integrate 1/pi int_0^inf dw a / (w^2 + a^2) = 1/2."""
    for w, p in (LT.legendre_frequency_grid(1.0, 20.0, 40, 40, 40), LT.regular_frequency_grid(5.0, 400, 400)):
        for a in (0.3, 1.0, 4.0):
            assert abs(np.dot(p, a / (w ** 2 + a ** 2)) - 0.5) < 2e-4


def test_kgrid_conventions():
    pts = LR.kgrid_points((2, (3, 2), 0.5))
    assert pts.shape == (6, 3)
    assert np.allclose(sorted(set(np.round(pts[:, 0], 12))), [1 / 6, 0.5, 5 / 6])
    assert np.allclose(sorted(set(np.round(pts[:, 1], 12))), [0.25, 0.75])
    assert np.allclose(LR.kgrid_points((0, (5,), 0.5)), [[0.0, 0.0, 0.0]])
