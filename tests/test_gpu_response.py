"""Frequency integrals and responses of an instance on the device
(Green_function_average, Green_function_density, tr_sigma_inf,
susceptibility, susceptibility_poles) against quantities fixed without the
Green-function path:
  * U = 0: the integrated Green function is the occupied one-body projector;
  * interacting: its elements are the ground-state averages <n_a>, <c^+_a c_b
    + h.c.> that cluster_averages measures by applying the operators to the
    ground state;
  * the self-energy at infinite frequency is the Hartree term, so its trace
    is U times the number of electrons on the interacting sites;
  * the response functions equal the Lehmann sums of the dense Jordan-Wigner
    Hamiltonian (tests/dense_ref.py), diagonalized with numpy.
Each check runs for the three Green-function methods where the method
matters (L Lehmann, F continued fractions, M matrix continued fractions)."""
import numpy as np
import pytest

from pyqcm_amd import _lib, qcm
from pyqcm_amd._lib import EDError
from tests.dense_ref import dense_hamiltonian

pytestmark = pytest.mark.gpu

_label = [9000]
_models = set()
METHODS = ["L", "F", "M"]


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    qcm.set_global_parameterC("GF_method", "L")
    yield
    qcm.set_global_parameterC("GF_method", "L")


def _chain_ops(L, extra=None):
    """Hubbard chain operators (0-based elements, both spins) plus probes."""
    ops = {
        "t": ("one-body", [(i, i + 1, -1.0) for i in range(L - 1)] + [(i + L, i + 1 + L, -1.0) for i in range(L - 1)]),
        "U": ("interaction", [(i, i + L, 1.0) for i in range(L)]),
        "mu": ("one-body", [(i, i, -1.0) for i in range(L)] + [(i + L, i + L, -1.0) for i in range(L)]),
        # probes (value 0 in the instances: they do not change H)
        "n0": ("one-body", [(0, 0, 1.0), (L, L, 1.0)]),
        "n0u": ("one-body", [(0, 0, 1.0)]),
        "n0d": ("one-body", [(L, L, 1.0)]),
        "h01": ("one-body", [(0, 1, 1.0), (L, L + 1, 1.0)]),
        "h01u": ("one-body", [(0, 1, 1.0)]),
        "h12u": ("one-body", [(1, 2, 1.0)]),
        "sz": ("one-body", [(i, i, 0.5 * (-1) ** i) for i in range(L)] + [(i + L, i + L, -0.5 * (-1) ** i) for i in range(L)]),
    }
    if extra:
        ops.update(extra)
    return ops


def _instance(name, n_sites, n_bath, ops, values, sector, method="L", solve=True):
    if name not in _models:
        qcm.new_model(name, n_sites, n_bath)
        for op, (typ, el) in ops.items():
            qcm.new_operator(name, op, typ, [(r + 1, c + 1, v) for (r, c, v) in el])
        _models.add(name)
    qcm.set_global_parameterC("GF_method", method)
    _label[0] += 1
    lab = _label[0]
    vals = dict(values)
    for op in ops:
        vals.setdefault(op, 0.0)
    qcm.new_model_instanceC(name, vals, sector, lab)
    if solve:
        qcm.ground_state_solve(lab)
    return lab


def _finish(lab):
    qcm.erase_model_instance(lab)
    qcm.set_global_parameterC("GF_method", "L")


@pytest.mark.parametrize("method", METHODS)
def test_average_u0_is_the_occupied_projector(method):
    L, t, mu = 4, 1.0, 0.35
    h = -mu * np.eye(L)
    for i in range(L - 1):
        h[i, i + 1] = h[i + 1, i] = -t
    eps, u = np.linalg.eigh(h)
    nneg = int(np.sum(eps < 0))
    lab = _instance("resp_chain4", L, 0, _chain_ops(L), {"t": t, "U": 0.0, "mu": mu}, "R0:N%d:S0" % (2 * nneg), method)
    P = u[:, eps < 0] @ u[:, eps < 0].T
    if method == "M":
        # only nneg of the hole vectors c_i |Omega> are independent; the block
        # Lanczos start (build_mcf, src_ed/model_instance.hpp:888-902) stops the
        # Gram-Schmidt at the first dependent one, c_nneg |Omega>, whose
        # projections it keeps, and drops the later columns of the weight
        P[nneg + 1:, :] = 0.0
        P[:, nneg + 1:] = 0.0
    M = qcm.Green_function_average(lab)
    assert M.shape == (L, L)
    assert np.abs(M - P).max() < 1e-10, (method, np.abs(M - P).max())
    assert abs(qcm.Green_function_density(lab) - 2.0 * np.trace(P) / L) < 1e-10
    _finish(lab)


@pytest.mark.parametrize("method", METHODS)
def test_average_interacting_matches_ground_state_averages(method):
    L, U = 4, 4.0
    lab = _instance("resp_chain4", L, 0, _chain_ops(L), {"t": 1.0, "U": U, "mu": U / 2}, "R0:N4:S0", method)
    ave = qcm.cluster_averages(lab)
    M = qcm.Green_function_average(lab)
    assert np.abs(M.imag).max() == 0.0
    assert np.abs(M - M.T).max() < 1e-12
    assert abs(2.0 * M[0, 0].real - ave["n0"][0]) < 1e-10, (method, M[0, 0], ave["n0"])
    assert abs(4.0 * M[0, 1].real - ave["h01"][0]) < 1e-10, (method, M[0, 1], ave["h01"])
    # half filling, particle-hole symmetric: one electron per site
    assert abs(qcm.Green_function_density(lab) - 1.0) < 1e-10
    _finish(lab)


@pytest.mark.parametrize("method", METHODS)
def test_average_two_degenerate_states(method):
    """3-site chain at half filling: the S = +1 and S = -1 doublet states,
    weight 1/2 each, spin blocks up and down.  The diagonal and, for L and M,
    the whole matrix are the weighted state averages.  With continued fractions
    the off-diagonal of the second state is assembled against the running sum,
    (b_0(o1, o2) - G_o1o1 - G_o2o2) / 2, so it carries minus half the first
    state's two diagonal elements (the reference's accumulation); the first
    state is the one of the first target sector, S = -1 (sectors ordered by
    S first, ascending: std::less<sector>, src_ed/sector.hpp:150-162)."""
    L, U = 3, 4.0
    ops = _chain_ops(L)
    vals = {"t": 1.0, "U": U, "mu": U / 2}
    lab = _instance("resp_chain3", L, 0, ops, vals, "R0:N3:S1/R0:N3:S-1", method)
    E, w = qcm.states(lab)
    assert len(E) == 2 and np.allclose(w, 0.5)
    ave = qcm.cluster_averages(lab)
    Mu, Md = qcm.Green_function_average(lab, 0), qcm.Green_function_average(lab, 1)
    assert abs(Mu[0, 0].real - ave["n0u"][0]) < 1e-10
    assert abs(Md[0, 0].real - ave["n0d"][0]) < 1e-10
    off = 2.0 * Mu[0, 1].real
    if method == "F":
        first = _instance("resp_chain3", L, 0, ops, vals, "R0:N3:S-1", "L")
        a1 = qcm.cluster_averages(first)
        Mf = qcm.Green_function_average(first, 0)
        _finish(first)
        quirk = -0.5 * 0.5 * (Mf[0, 0].real + Mf[1, 1].real)
        assert abs(off - (ave["h01u"][0] + 2.0 * quirk)) < 1e-10, (off, ave["h01u"], quirk)
        assert abs(Mf[0, 0].real - a1["n0u"][0]) < 1e-10
    else:
        assert abs(off - ave["h01u"][0]) < 1e-10, (method, off, ave["h01u"])
        assert abs(2.0 * Mu[1, 2].real - ave["h12u"][0]) < 1e-10
    assert abs(qcm.Green_function_density(lab) - 1.0) < 1e-10
    _finish(lab)


def test_average_after_batched_solve():
    """continued fractions held in the batched solver's device store: the
    same averages as the per-instance solve, and as the operator averages"""
    L = 4
    points = [(4.0, 2.0), (2.5, 1.25), (6.0, 2.7)]
    labs = [_instance("resp_chain4", L, 0, _chain_ops(L), {"t": 1.0, "U": U, "mu": mu}, "R0:N4:S0", "F", solve=False)
            for (U, mu) in points]
    qcm.set_global_parameterC("GF_method", "F")
    qcm.solve_batch(labs)
    for lab, (U, mu) in zip(labs, points):
        M = qcm.Green_function_average(lab)
        one = _instance("resp_chain4", L, 0, _chain_ops(L), {"t": 1.0, "U": U, "mu": mu}, "R0:N4:S0", "F")
        ref = qcm.Green_function_average(one)
        ave = qcm.cluster_averages(one)
        _finish(one)
        assert np.abs(M - ref).max() < 1e-10, (U, np.abs(M - ref).max())
        assert abs(4.0 * M[0, 1].real - ave["h01"][0]) < 1e-10
    for lab in labs:
        _finish(lab)


@pytest.mark.parametrize("method", METHODS)
def test_tr_sigma_inf_is_the_hartree_term(method):
    """Sigma(infinity)_{i sigma} = U <n_{i, -sigma}>: the trace over sites and
    spins is U <N> on the interacting sites.  Chain (4 electrons, normal),
    3-site doublet (up/down blocks) and a 2 + 2 bath model."""
    lab = _instance("resp_chain4", 4, 0, _chain_ops(4), {"t": 1.0, "U": 4.0, "mu": 2.0}, "R0:N4:S0", method)
    assert abs(qcm.tr_sigma_inf(lab) - 16.0) < 1e-6, (method, qcm.tr_sigma_inf(lab))
    _finish(lab)
    lab = _instance("resp_chain3", 3, 0, _chain_ops(3), {"t": 1.0, "U": 3.0, "mu": 1.5}, "R0:N3:S1/R0:N3:S-1", method)
    assert abs(qcm.tr_sigma_inf(lab) - 9.0) < 1e-6, (method, qcm.tr_sigma_inf(lab))
    _finish(lab)
    ns, nb = 2, 2
    Lb = ns + nb
    one = {"t": [(0, 1, -1.0)], "V": [(0, 2, 0.5), (1, 3, 0.4)], "eb": [(2, 2, 1.0), (3, 3, -0.7)],
           "mu": [(0, 0, -1.0), (1, 1, -1.0)], "nc": [(0, 0, 1.0), (1, 1, 1.0)]}
    ops = {k: ("one-body", el + [(r + Lb, c + Lb, v) for (r, c, v) in el]) for k, el in one.items()}
    ops["U"] = ("interaction", [(i, i + Lb, 1.0) for i in range(ns)])
    U = 2.5
    lab = _instance("resp_bath", ns, nb, ops, {"t": 1.0, "V": 1.0, "eb": 1.0, "mu": 1.0, "U": U}, "R0:N4:S0", method)
    nc = qcm.cluster_averages(lab)["nc"][0]
    assert abs(qcm.tr_sigma_inf(lab) - U * nc) < 1e-6, (method, qcm.tr_sigma_inf(lab), U * nc)
    _finish(lab)


def _dense_lehmann(L, ops, vals, op, Nup, Ndw):
    """poles E_n - E_0 and residues |<n|O|0>|^2 of the sector's ground state."""
    terms = [(typ, float(vals.get(name, 0.0)), el) for name, (typ, el) in ops.items()]
    H = dense_hamiltonian(L, terms, Nup, Ndw)
    E, U = np.linalg.eigh(H)
    O = dense_hamiltonian(L, [(ops[op][0], 1.0, ops[op][1])], Nup, Ndw)
    r = (U.T @ (O @ U[:, 0])) ** 2
    return E[0], E - E[0], r


def _merge(poles, res, weight):
    """poles above 1e-8 with residue weight r > 1e-6, merged within 1e-4"""
    out = []
    for e, r in zip(poles, res * weight):
        if e < 1e-8 or r <= 1e-6:
            continue
        for k in range(len(out)):
            if abs(out[k][0] - e) < 1e-4:
                out[k][1] += r
                break
        else:
            out.append([e, r])
    return np.array(sorted(out)).reshape(-1, 2)


@pytest.mark.parametrize("op", ["n0", "sz", "h01", "U"])
def test_susceptibility_matches_dense_lehmann(op):
    L, U = 4, 4.0
    ops = _chain_ops(L)
    vals = {"t": 1.0, "U": U, "mu": U / 2}
    lab = _instance("resp_chain4", L, 0, ops, vals, "R0:N4:S0")
    _, poles, res = _dense_lehmann(L, ops, vals, op, 2, 2)
    w = np.array([0.3 + 0.1j, 1.7 + 0.05j, -0.5 + 0.2j, 2.0j, 6.5 + 0.3j])
    chi = qcm.susceptibility(op, w, lab)
    ref = np.array([np.sum(res / (x - poles)) for x in w])
    assert np.abs(chi - ref).max() < 1e-9 * max(1.0, np.abs(ref).max()), (op, np.abs(chi - ref).max())
    got = qcm.susceptibility_poles(op, lab)
    exp = _merge(poles, res, 1.0)
    got = got[np.argsort(got[:, 0])]
    assert got.shape == exp.shape, (op, got, exp)
    assert np.abs(got - exp).max() < 1e-8, (op, got, exp)
    _finish(lab)


def test_susceptibility_two_states():
    """the weighted sum over the S = +1 and S = -1 states of the 3-site chain"""
    L, U = 3, 4.0
    ops = _chain_ops(L)
    vals = {"t": 1.0, "U": U, "mu": U / 2}
    lab = _instance("resp_chain3", L, 0, ops, vals, "R0:N3:S1/R0:N3:S-1")
    w = np.array([0.4 + 0.1j, 2.5 + 0.2j, 1.0j])
    for op in ("n0u", "sz"):
        ref = np.zeros(len(w), dtype=complex)
        pol = []
        for Nup, Ndw in ((2, 1), (1, 2)):
            _, poles, res = _dense_lehmann(L, ops, vals, op, Nup, Ndw)
            ref += 0.5 * np.array([np.sum(res / (x - poles)) for x in w])
            pol.append((poles, res))
        chi = qcm.susceptibility(op, w, lab)
        assert np.abs(chi - ref).max() < 1e-9, (op, np.abs(chi - ref).max())
        # residue sum of the kept poles: every excitation with weight above
        # the 1e-6 threshold, whichever state it comes from
        got = qcm.susceptibility_poles(op, lab)
        tot = sum(np.sum(0.5 * r[(p >= 1e-8) & (0.5 * r > 1e-6)]) for p, r in pol)
        assert abs(got[:, 1].sum() - tot) < 1e-8, (op, got, tot)
    _finish(lab)


def test_response_errors():
    L = 4
    lab = _instance("resp_chain4", L, 0, _chain_ops(L), {"t": 1.0, "U": 4.0, "mu": 2.0}, "R0:N4:S0")
    with pytest.raises(EDError) as e:
        qcm.susceptibility("nope", [0.5j], lab)
    assert e.value.code == 1
    qcm.new_model("resp_chain4b", L, 0)
    for op, (typ, el) in _chain_ops(L).items():
        qcm.new_operator("resp_chain4b", op, typ, [(r + 1, c + 1, v) for (r, c, v) in el])
    _label[0] += 1
    lab2 = _label[0]
    qcm.new_model_instanceC("resp_chain4b", {"t": 1.0, "U": 4.0, "mu": 2.0}, "R0:N4:S0", lab2)
    with pytest.raises(EDError) as e:  # defined in the model, not active in the instance
        qcm.susceptibility_poles("sz", lab2)
    assert e.value.code == 1
    assert qcm.susceptibility("n0", [], lab).shape == (0,)
    qcm.erase_model_instance(lab2)
    _finish(lab)


def test_interrupt_stops_the_next_step_once():
    """a pending interrupt stops the ground-state and Green-function solves at
    their first step with code 11 and is consumed: the calls then succeed"""
    import threading
    import time
    L = 4
    lab = _instance("resp_chain4", L, 0, _chain_ops(L), {"t": 1.0, "U": 4.0, "mu": 2.0}, "R0:N4:S0", solve=False)
    qcm.interrupt()
    with pytest.raises(EDError) as e:
        qcm.ground_state_solve(lab)
    assert e.value.code == 11
    assert str(e.value).startswith("[cancelled]"), str(e.value)
    E0, _ = qcm.ground_state_solve(lab)
    qcm.interrupt()
    with pytest.raises(EDError) as e:
        qcm.Green_function_solveC(lab)
    assert e.value.code == 11
    G = qcm.Green_function(0.5j, label=lab)
    assert np.all(np.isfinite(G))
    _finish(lab)
    # from another thread, during a longer solve (Davidson on 3 states of a
    # 12-orbital sector, ~70 000 states)
    qcm.new_model("resp_chain12", 12, 0)
    for op, (typ, el) in _chain_ops(12).items():
        qcm.new_operator("resp_chain12", op, typ, [(r + 1, c + 1, v) for (r, c, v) in el])
    _label[0] += 1
    lab = _label[0]
    qcm.new_model_instanceC("resp_chain12", {"t": 1.0, "U": 4.0, "mu": 2.0}, "R0:N12:S0", lab)
    H = qcm.Hamiltonian(lab, "R0:N12:S0", "G")
    out = {}

    def run():
        t0 = time.perf_counter()
        try:
            H.davidson(3)
            out["code"] = 0
        except EDError as err:
            out["code"] = err.code
        out["s"] = time.perf_counter() - t0

    th = threading.Thread(target=run)
    th.start()
    time.sleep(0.05)
    qcm.interrupt()
    th.join(60)
    assert not th.is_alive()
    # either the request landed inside the iteration (code 11) or the solve
    # had already finished; a request left pending is consumed by the next step
    assert out["code"] in (0, 11), out
    if out["code"] == 0:
        with pytest.raises(EDError):
            H.davidson(1)
    del H
    qcm.erase_model_instance(lab)


def test_concurrent_callers_get_the_serial_results():
    """four threads solving and evaluating different instances at once (the
    reference's OpenMP loop over clusters) get the results of serial calls"""
    import threading
    L = 4
    pts = [(3.0, 1.5), (4.0, 2.0), (5.0, 2.2), (6.0, 3.0)]
    z = np.array([0.3 + 0.1j, 1.1j, -0.7 + 0.2j])
    labs = [_instance("resp_chain4", L, 0, _chain_ops(L), {"t": 1.0, "U": U, "mu": mu}, "R0:N4:S0", solve=False)
            for (U, mu) in pts]
    got = {}

    def run(lab):
        E0, _ = qcm.ground_state_solve(lab)
        got[lab] = (E0, qcm.Green_function_grid(z, label=lab), qcm.Green_function_average(lab))

    ths = [threading.Thread(target=run, args=(lab,)) for lab in labs]
    for t in ths:
        t.start()
    for t in ths:
        t.join(120)
    assert all(not t.is_alive() for t in ths) and len(got) == len(labs)
    for lab, (U, mu) in zip(labs, pts):
        one = _instance("resp_chain4", L, 0, _chain_ops(L), {"t": 1.0, "U": U, "mu": mu}, "R0:N4:S0")
        E0 = qcm.ground_state_solve(one)[0]
        G = qcm.Green_function_grid(z, label=one)
        M = qcm.Green_function_average(one)
        _finish(one)
        assert abs(got[lab][0] - E0) <= 1e-12 * abs(E0)
        assert np.abs(got[lab][1] - G).max() < 1e-10
        assert np.abs(got[lab][2] - M).max() < 1e-10
    for lab in labs:
        _finish(lab)


def test_fidelity_matches_dense_overlaps():
    """|<psi(U=4)|psi(U=4.6)>|^2 against the dense ground states; an instance
    with itself gives the sum of the squared weights of its states"""
    L = 4
    ops = _chain_ops(L)
    a = _instance("resp_chain4", L, 0, ops, {"t": 1.0, "U": 4.0, "mu": 2.0}, "R0:N4:S0")
    b = _instance("resp_chain4", L, 0, ops, {"t": 1.0, "U": 4.6, "mu": 2.3}, "R0:N4:S0")
    gs = []
    for U, mu in ((4.0, 2.0), (4.6, 2.3)):
        vals = {"t": 1.0, "U": U, "mu": mu}
        terms = [(typ, float(vals.get(name, 0.0)), el) for name, (typ, el) in ops.items()]
        E, V = np.linalg.eigh(dense_hamiltonian(L, terms, 2, 2))
        assert E[1] - E[0] > 1e-3
        gs.append(V[:, 0])
    ref = float(gs[0] @ gs[1]) ** 2
    assert abs(qcm.fidelity(a, b) - ref) < 1e-10, (qcm.fidelity(a, b), ref)
    assert abs(qcm.fidelity(a, a) - 1.0) < 1e-12
    d = _instance("resp_chain3", 3, 0, _chain_ops(3), {"t": 1.0, "U": 4.0, "mu": 2.0}, "R0:N3:S1/R0:N3:S-1")
    assert abs(qcm.fidelity(d, d) - 0.5) < 1e-12
    assert qcm.fidelity(a, d) == 0.0
    # same sector name, different model size: refused (the vectors differ in length)
    f = _instance("resp_chain5", 5, 0, _chain_ops(5), {"t": 1.0, "U": 4.0, "mu": 2.0}, "R0:N4:S0")
    with pytest.raises(EDError) as e:
        qcm.fidelity(a, f)
    assert e.value.code == 1
    for x in (a, b, d, f):
        _finish(x)


def test_density_matrix():
    """reduced density matrices of the 4-site chain ground state: one site
    (diagonal, the occupation probabilities of the dense ground state), all
    sites (|psi><psi| in label order), and two complementary halves (the same
    nonzero spectrum, as for any pure state)"""
    L = 4
    ops = _chain_ops(L)
    vals = {"t": 1.0, "U": 4.0, "mu": 2.0}
    lab = _instance("resp_chain4", L, 0, ops, vals, "R0:N4:S0")
    terms = [(typ, float(vals.get(name, 0.0)), el) for name, (typ, el) in ops.items()]
    E, V = np.linalg.eigh(dense_hamiltonian(L, terms, 2, 2))
    psi = V[:, 0]
    from tests.dense_ref import sector_states
    words = sector_states(L, 2, 2)  # label order, word = up | down << L
    rho, basis = qcm.density_matrix([0], lab)
    assert sorted(int(b) for b in basis) == [0, 1, 1 << 32, 1 + (1 << 32)]
    p = {0: 0.0, 1: 0.0, 1 << 32: 0.0, 1 + (1 << 32): 0.0}
    for k, w in enumerate(words):
        key = (w & 1) | (((w >> L) & 1) << 32)
        p[key] += psi[k] ** 2
    assert np.abs(rho - np.diag(np.diag(rho))).max() < 1e-12
    for i, b in enumerate(basis):
        assert abs(rho[i, i].real - p[int(b)]) < 1e-10
    rho_all, basis_all = qcm.density_matrix(list(range(L)), lab)
    assert len(basis_all) == len(words)
    # all orbitals: the subsystem states are the basis states in label order
    assert [int(b) for b in basis_all] == [(w & 0xF) | ((w >> L) << 32) for w in words]
    assert np.abs(rho_all.real - np.outer(psi, psi)).max() < 1e-10
    ra, _ = qcm.density_matrix([0, 1], lab)
    rb, _ = qcm.density_matrix([2, 3], lab)
    for r in (ra, rb):
        assert np.abs(r - r.T.conj()).max() < 1e-12 and abs(np.trace(r).real - 1.0) < 1e-12
    ea = np.sort(np.linalg.eigvalsh(ra.real))[::-1]
    eb = np.sort(np.linalg.eigvalsh(rb.real))[::-1]
    k = min(len(ea), len(eb))
    assert np.abs(ea[:k] - eb[:k]).max() < 1e-10 and np.abs(ea[k:]).max(initial=0.0) < 1e-10
    with pytest.raises(EDError):
        qcm.density_matrix([7], lab)
    _finish(lab)


def test_print_wavefunction():
    """the text form of the ground state: header (sector, energy, weight),
    then, for sectors up to max_dim_print states, |psi|^2, psi and the basis
    state |up bits,down bits> (least significant orbital first) in label order;
    the probabilities are those of the dense ground state (6 printed digits)"""
    L = 4
    ops = _chain_ops(L)
    vals = {"t": 1.0, "U": 4.0, "mu": 2.0}
    lab = _instance("resp_chain4", L, 0, ops, vals, "R0:N4:S0")
    txt = qcm.print_wavefunction(lab)
    lines = txt.rstrip("\n").split("\n")
    assert lines[0] == "state"
    sec, e, w = lines[1].split("\t")
    assert sec == "R0:N4:S0" and abs(float(e) - (-9.9531453)) < 1e-4 and float(w) == 1.0
    from tests.dense_ref import sector_states
    words = sector_states(L, 2, 2)
    assert len(lines) == 2 + len(words)
    terms = [(typ, float(vals.get(name, 0.0)), el) for name, (typ, el) in ops.items()]
    _, V = np.linalg.eigh(dense_hamiltonian(L, terms, 2, 2))
    for k, (line, word) in enumerate(zip(lines[2:], words)):
        p2, amp, ket = line.split("\t")
        up, dn = word & 0xF, word >> L
        assert ket == "|%s,%s>" % ("".join("1" if up >> i & 1 else "0" for i in range(L)),
                                   "".join("1" if dn >> i & 1 else "0" for i in range(L)))
        assert abs(float(p2) - V[k, 0] ** 2) < 1e-5 * max(1.0, V[k, 0] ** 2) + 1e-7
        assert abs(abs(float(amp)) - abs(V[k, 0])) < 1e-5
    qcm.set_global_parameterC("max_dim_print", 10)
    try:
        assert qcm.print_wavefunction(lab).rstrip("\n").split("\n") == lines[:2]
    finally:
        qcm.set_global_parameterC("max_dim_print", 64)
    _finish(lab)
