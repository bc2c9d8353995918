"""Lattice sums on the device (ed_cpt_projected_gf, ed_sef_grid) against the
numpy oracle on the same cluster Green functions; the non-interacting limit
end to end (CPT is the lattice Green function, the grid Potthoff functional
is the lattice grand potential); and the cluster self-energy, hybridization
function and Potthoff functional of a bath model against closed forms.
Tolerances: 1e-10 (device vs oracle on the same G), 1e-9 (CPT closed form
through the whole ED path), 1e-8 (Potthoff functional, limited by the
frequency quadrature)."""
import numpy as np
import pytest

from oracle import lattice_ref as LR
from pyqcm_amd import _lib, lattice as LT, qcm, workloads as W

pytestmark = pytest.mark.gpu

_label = [7000]
_models = set()


def _new_label():
    _label[0] += 1
    return _label[0]


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    qcm.set_global_parameterC("GF_method", "L")
    yield


def _instance(name, n_sites, n_bath, operators, values, sector):
    if name not in _models:
        qcm.new_model(name, n_sites, n_bath)
        for op, (typ, el) in operators.items():
            qcm.new_operator(name, op, typ, [(r + 1, c + 1, v) for (r, c, v) in el])
        _models.add(name)
    lab = _new_label()
    qcm.new_model_instanceC(name, values, sector, lab)
    qcm.ground_state_solve(lab)
    return lab


def _natural(G):
    """blocks as the binding exposes them (the reference's flat column-major
    copy, src_python/qcm_nb.hpp:27-41) -> G[iz, a, b] = G_ab"""
    return np.transpose(np.asarray(G), (0, 2, 1))


def _neel_cluster(M=0.15):
    wl = W.vca_cluster(2, 2)
    lab = _instance("lat_" + wl.name, wl.n_sites, wl.n_bath, wl.operators, {"t": 1.0, "U": 4.0, "mu": 0.4, "M": M},
                    wl.sector)
    s = np.array([1.0 if (x + y) % 2 == 0 else -1.0 for y in range(2) for x in range(2)])
    return lab, LT.square_superlattice(2, 2, 1.0, weiss=M * s), LT.square_superlattice(2, 2, 1.0, weiss=-M * s)


def test_cpt_projected_gf_matches_oracle_on_a_neel_cluster():
    lab, sl_up, sl_dn = _neel_cluster()
    z = np.concatenate([np.linspace(-4.0, 4.0, 9) + 0.2j, 1j * np.array([0.1, 0.5, 1.3, 4.0])])
    kg = (2, (8, 8), 0.5)
    for spin, sl in ((0, sl_up), (1, sl_dn)):
        got = _natural(qcm.projected_Green_function(z, sl, kg, spin_down=spin, label=lab))
        ref = LR.cpt_projected_gf(_natural(qcm.Green_function_grid(z, spin_down=spin, label=lab)), sl, kg)
        assert np.abs(got - ref).max() < 1e-10 * max(1.0, np.abs(ref).max()), (spin, np.abs(got - ref).max())
    qcm.erase_model_instance(lab)


def test_sef_grid_matches_oracle_on_a_neel_cluster():
    lab, sl_up, sl_dn = _neel_cluster()
    w, p = LT.legendre_frequency_grid(1.0, 20.0, 16, 16, 16)
    kg = (2, (6, 6), 0.0)
    I, per = qcm.SEF_grid(w, p, sl_up, kg, sl_dn, label=lab)
    assert qcm.mixingC(lab) & 4
    ref = np.zeros(len(w))
    for spin, sl in ((0, sl_up), (1, sl_dn)):
        G = _natural(qcm.Green_function_grid(1j * w, spin_down=spin, label=lab))
        ref += LR.sef_logdet(G, sl, kg)
    assert np.abs(per - ref).max() < 1e-10, np.abs(per - ref).max()
    assert abs(I - np.dot(p, ref)) < 1e-10 * max(1.0, abs(I))
    qcm.erase_model_instance(lab)


@pytest.mark.parametrize("L", [3, 4, 5, 6])
def test_non_interacting_chain_end_to_end(L):
    """U = 0: CPT is exact, and the grid Potthoff functional is the grand
    potential of the lattice (also with a Weiss field on the cluster only).
    L = 3, 4 / 5 / 6 put 4, 4 / 2 / 1 matrices in one wave."""
    t, mu, nk = 1.0, 0.35, 24
    # the shift keeps every cluster level >= 0.095 away from zero (a zero level
    # is a degenerate ground state and a pole on the integration axis)
    for weiss in (None, 0.25 * np.array([(-1.0) ** i for i in range(L)]) + 0.22):
        ops = dict(W.hubbard_chain(L, U=0.0, mu=mu).operators)
        vals = {"t": t, "U": 0.0, "mu": mu}
        if weiss is not None:
            ops["w"] = ("one-body", [(i, i, weiss[i]) for i in range(L)] + [(i + L, i + L, weiss[i]) for i in range(L)])
            vals["w"] = 1.0
        name = "lat_chain%d_u0" % L + ("_w" if weiss is not None else "")
        # grand-canonical ground state: as many electrons as negative one-body levels
        hc = np.diag(-mu * np.ones(L) + (weiss if weiss is not None else 0.0))
        for i in range(L - 1):
            hc[i, i + 1] = hc[i + 1, i] = -t
        nneg = int(np.sum(np.linalg.eigvalsh(hc) < 0))
        lab = _instance(name, L, 0, ops, vals, "R0:N%d:S0" % (2 * nneg))
        sl = LT.chain_superlattice(L, t, weiss=weiss)
        # CPT average = local lattice Green function (Bloch sum)
        z = np.array([-1.1 + 0.1j, 0.3 + 0.05j, 2.2 + 0.2j, 0.7j])
        P = _natural(qcm.projected_Green_function(z, sl, (1, (nk,), 0.5), label=lab))
        ks = (np.arange(nk) + 0.5) / nk
        eps = np.array([[-2 * t * np.cos(2 * np.pi * (k + m) / L) - mu for m in range(L)] for k in ks])
        for iz, zz in enumerate(z):
            assert abs(np.trace(P[iz]) / L - np.mean(1.0 / (zz - eps))) < 1e-9
        # Potthoff functional on the grid = lattice grand potential per site
        w, p = LT.legendre_frequency_grid(1.0, 20.0, 60, 60, 60)
        omega = LT.potthoff_functional(lab, w, p, sl, (1, (nk,), 0.0))
        eps0 = np.array([[-2 * t * np.cos(2 * np.pi * (k + m) / L) - mu for m in range(L)] for k in np.arange(nk) / nk])
        # bound set by the frequency quadrature (a level near zero makes the
        # integrand steep at small omega), not by the k-sums
        assert abs(omega - 2.0 * np.sum(np.minimum(eps0, 0.0)) / (nk * L)) < 1e-8, (weiss, omega)
        qcm.erase_model_instance(lab)


def test_bath_model_self_energy_hybridization_and_potthoff():
    """2 sites + 2 bath orbitals with cluster, cluster-bath and bath-bath
    hoppings: at U = 0 the self-energy vanishes; the hybridization function is
    T_cb (z - T_bb)^{-1} T_bc; the cluster Potthoff functional is E0 minus
    twice the negative bath levels."""
    ns, nb = 2, 2
    L = ns + nb
    one = {"t": [(0, 1, -1.0)], "V": [(0, 2, 0.5), (1, 3, 0.4), (0, 3, 0.2)], "tb": [(2, 3, 0.3)],
           "eb": [(2, 2, 1.0), (3, 3, -0.7)], "mu": [(0, 0, -1.0), (1, 1, -1.0)]}
    ops = {k: ("one-body", el + [(r + L, c + L, v) for (r, c, v) in el]) for k, el in one.items()}
    ops["U"] = ("interaction", [(i, i + L, 1.0) for i in range(ns)])
    vals = {"t": 1.0, "V": 1.0, "tb": 1.0, "eb": 1.0, "mu": 0.2, "U": 0.0}
    h = np.zeros((L, L))
    for k, el in one.items():
        for r, c, v in el:
            h[r, c] += vals[k] * v
            if r != c:
                h[c, r] += vals[k] * v
    nneg = int(np.sum(np.linalg.eigvalsh(h) < 0))
    lab = _instance("lat_bath", ns, nb, ops, vals, "R0:N%d:S0" % (2 * nneg))
    Tcb, Tbb = h[:ns, ns:], h[ns:, ns:]
    for zz in (0.3 + 0.1j, -1.2 + 0.05j, 0.8j, 2.5 + 0.4j):
        S = np.asarray(qcm.self_energyC(zz, label=lab))
        assert np.abs(S).max() < 1e-9, (zz, np.abs(S).max())
        Gam = np.asarray(qcm.hybridization_functionC(zz, label=lab)).T
        ref = Tcb @ np.linalg.inv(zz * np.eye(nb) - Tbb) @ Tcb.T
        assert np.abs(Gam - ref).max() < 1e-12, (zz, np.abs(Gam - ref).max())
    E0 = qcm.ground_state_solve(lab)[0]
    eb = np.linalg.eigvalsh(Tbb)
    assert abs(qcm.Potthoff_functionalC(lab) - (E0 - 2.0 * np.sum(eb[eb < 0]))) < 1e-10
    qcm.erase_model_instance(lab)


def test_lattice_sums_reject_bad_input():
    """Errors come back as EDError with the library's status codes: a
    superlattice of the wrong dimension or with an element outside it, a
    missing spin-down superlattice under spin mixing, and d > 8."""
    from pyqcm_amd._lib import EDError
    lab, up, dn = _neel_cluster()
    kg = (2, (4, 4), 0.5)
    bad_dim = LT.square_superlattice(2, 1, 1.0)
    with pytest.raises(EDError) as e:
        qcm.projected_Green_function([0.5j], bad_dim, kg, label=lab)
    assert e.value.code == 1
    bad_el = dict(up)
    bad_el["elements"] = list(up["elements"]) + [(0, 7, 0, -1.0)]
    with pytest.raises(EDError) as e:
        qcm.projected_Green_function([0.5j], bad_el, kg, label=lab)
    assert e.value.code == 1
    w, p = LT.legendre_frequency_grid(1.0, 20.0, 4, 4, 4)
    with pytest.raises(EDError) as e:
        qcm.SEF_grid(w, p, up, kg, None, label=lab)
    assert e.value.code == 1
    qcm.erase_model_instance(lab)
    wl = W.hubbard_chain(9, U=4.0, mu=2.0, N=8)
    big = _instance("lat_chain9", wl.n_sites, wl.n_bath, wl.operators, dict(wl.values), wl.sector)
    with pytest.raises(EDError) as e:
        qcm.projected_Green_function([0.5j], LT.chain_superlattice(9), (1, (4,), 0.5), label=big)
    assert e.value.code == 8
    qcm.erase_model_instance(big)


@pytest.mark.parametrize("cdw,omega_ref,e0_ref", [(1e-9, -2.5525862, -9.9531453), (0.01, -2.5526019, -9.9531709),
                                                   (-0.01, -2.5526019, -9.9531709), (0.0, -2.5525862, -9.9531453)])
def test_potthoff_functional_reproduces_the_reference_1D_4_example(cdw, omega_ref, e0_ref):
    """The reference's own output (examples/1D_4/sef.tsv, first block): the
    Hubbard chain (t = 1, U = 4, mu = 2) tiled by 4-site clusters carrying a
    charge-density-wave Weiss field cdw_1 (site energies cdw_1 (-1)^i, both
    spins), sector N = 4, S = 0; columns omega and E0_1, printed to 8 digits.
    The other operators of those rows (V, V0m, V1m at 1e-9) change omega by
    less than the last printed digit and are left out."""
    L = 4
    ops = dict(W.hubbard_chain(L, U=4.0, mu=2.0).operators)
    stag = np.array([(-1.0) ** i for i in range(L)])
    ops["cdw"] = ("one-body", [(i, i, stag[i]) for i in range(L)] + [(i + L, i + L, stag[i]) for i in range(L)])
    lab = _instance("lat_chain4_cdw", L, 0, ops, {"t": 1.0, "U": 4.0, "mu": 2.0, "cdw": cdw}, "R0:N4:S0")
    E0 = qcm.ground_state_solve(lab)[0]
    assert abs(E0 - e0_ref) < 1e-7, (E0, e0_ref)
    w, p = LT.legendre_frequency_grid(1.0, 20.0, 60, 60, 60)
    omega = LT.potthoff_functional(lab, w, p, LT.chain_superlattice(L, 1.0, weiss=cdw * stag), (1, (128,), 0.0))
    assert abs(omega - omega_ref) < 2e-7, (cdw, omega, omega_ref)
    qcm.erase_model_instance(lab)
