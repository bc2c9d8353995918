"""Smallest sectors through the whole device path, against closed forms:
  * the Hubbard atom (one site, U, mu = U/2) in its S = +-1 doublet: E0 = -U/2
    and G(z) = 1/2 / (z + U/2) + 1/2 / (z - U/2) for every Green-function
    method, integrated G = 1/2, Sigma(inf) = U <N>;
  * the empty sector (one state, no hole part): E0 = 0 and G(z) = (z - h)^-1
    for the one-body matrix h of a 3-site chain with a bath-free hopping;
  * the fully filled sector (one state, no electron part): E0 and
    G(z) = (z - h - U)^-1 for on-site U, since removing an electron from the
    full band leaves the other spin's U on every site.
One-state sectors make one-step Krylov spaces (and, at zero energy, a zero
tridiagonal matrix); 14 p log(dim) steps, the reference's count, is 0 there,
so at least p + 1 band-Lanczos steps and one block floor are taken."""
import numpy as np
import pytest

from pyqcm_amd import _lib, qcm

pytestmark = pytest.mark.gpu

_label = [15000]
_models = set()


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    yield
    qcm.set_global_parameterC("GF_method", "L")


def _chain(L):
    name = "edge_chain%d" % L
    if name not in _models:
        qcm.new_model(name, L, 0)
        el = [(i, i + 1, -1.0) for i in range(L - 1)]
        qcm.new_operator(name, "t", "one-body", [(r + 1, c + 1, v) for (r, c, v) in el + [(r + L, c + L, v) for (r, c, v) in el]])
        qcm.new_operator(name, "U", "interaction", [(i + 1, i + L + 1, 1.0) for i in range(L)])
        qcm.new_operator(name, "mu", "one-body", [(i + 1, i + 1, -1.0) for i in range(L)] + [(i + L + 1, i + L + 1, -1.0) for i in range(L)])
        _models.add(name)
    return name


def _instance(L, vals, sector, method):
    qcm.set_global_parameterC("GF_method", method)
    _label[0] += 1
    qcm.new_model_instanceC(_chain(L), vals, sector, _label[0])
    return _label[0]


Z = np.array([0.3 + 0.2j, -1.7 + 0.1j, 2.5j, 4.0 + 0.05j])


@pytest.mark.parametrize("method", ["L", "F", "M"])
def test_hubbard_atom_doublet(method):
    U = 3.0
    lab = _instance(1, {"U": U, "mu": U / 2}, "R0:N1:S1/R0:N1:S-1", method)
    E0, _ = qcm.ground_state_solve(lab)
    assert abs(E0 + U / 2) < 1e-12
    for sd in (0, 1):
        G = qcm.Green_function_grid(Z, spin_down=sd, label=lab)[:, 0, 0]
        ref = 0.5 / (Z + U / 2) + 0.5 / (Z - U / 2)
        assert np.abs(G - ref).max() < 1e-12, (method, sd, np.abs(G - ref).max())
        assert abs(qcm.Green_function_average(lab, sd)[0, 0] - 0.5) < 1e-12
    assert abs(qcm.tr_sigma_inf(lab) - U * 1.0) < 1e-6
    qcm.erase_model_instance(lab)


def _h(L, t, mu):
    h = -mu * np.eye(L)
    for i in range(L - 1):
        h[i, i + 1] = h[i + 1, i] = -t
    return h


@pytest.mark.parametrize("method", ["L", "F", "M"])
def test_empty_sector(method):
    # every one-body level above zero, so that the empty state is the ground
    # state (the Lehmann average counts the poles below zero)
    L, t, mu, U = 3, 1.0, -2.0, 5.0
    lab = _instance(L, {"t": t, "U": U, "mu": mu}, "R0:N0:S0", method)
    E0, _ = qcm.ground_state_solve(lab)
    assert abs(E0) < 1e-14
    G = qcm.Green_function_grid(Z, label=lab)
    h = _h(L, t, mu)
    for iz, z in enumerate(Z):
        ref = np.linalg.inv(z * np.eye(L) - h)
        assert np.abs(G[iz] - ref).max() < 1e-10, (method, np.abs(G[iz] - ref).max())
    assert np.abs(qcm.Green_function_average(lab)).max() < 1e-14
    qcm.erase_model_instance(lab)


@pytest.mark.parametrize("method", ["L", "F", "M"])
def test_full_sector(method):
    # mu large enough that the removal energies lie below zero, as for a
    # ground state: the Lehmann average counts the poles below zero
    # (Q_matrix::integrated_Green_function), the fractions count all hole weight
    L, t, mu, U = 3, 1.0, 5.0, 2.5
    lab = _instance(L, {"t": t, "U": U, "mu": mu}, "R0:N6:S0", method)
    E0, _ = qcm.ground_state_solve(lab)
    h = _h(L, t, mu)
    # the full state: every one-body diagonal twice, U on every site
    assert abs(E0 - (2.0 * np.trace(h) + U * L)) < 1e-12
    G = qcm.Green_function_grid(Z, label=lab)
    for iz, z in enumerate(Z):
        ref = np.linalg.inv(z * np.eye(L) - h - U * np.eye(L))
        assert np.abs(G[iz] - ref).max() < 1e-10, (method, np.abs(G[iz] - ref).max())
    M = qcm.Green_function_average(lab)
    assert np.abs(M - np.eye(L)).max() < 1e-12
    qcm.erase_model_instance(lab)


def test_products_on_the_smallest_sectors():
    """H.x on sectors with a one-state half basis (no electron or a full band
    of one spin) and on the one-site model, every format, against the dense
    Jordan-Wigner Hamiltonian"""
    from tests.dense_ref import dense_hamiltonian
    for L, sectors in ((3, ["R0:N0:S0", "R0:N1:S1", "R0:N1:S-1", "R0:N2:S2", "R0:N3:S3", "R0:N5:S1", "R0:N6:S0"]),
                       (1, ["R0:N0:S0", "R0:N1:S1", "R0:N2:S0"])):
        vals = {"t": 1.0, "U": 2.5, "mu": 0.7} if L > 1 else {"U": 2.5, "mu": 0.7}
        lab = _instance(L, vals, sectors[0], "L")
        el_t = [(i, i + 1, -1.0) for i in range(L - 1)]
        terms = []
        if L > 1:
            terms.append(("one-body", 1.0, el_t + [(r + L, c + L, v) for (r, c, v) in el_t]))
        terms.append(("interaction", 2.5, [(i, i + L, 1.0) for i in range(L)]))
        terms.append(("one-body", 0.7, [(i, i, -1.0) for i in range(L)] + [(i + L, i + L, -1.0) for i in range(L)]))
        for sec in sectors:
            N, S = [int(x[1:]) for x in sec.split(":")[1:]]
            Nup, Ndw = (N + S) // 2, (N - S) // 2
            Hd = dense_hamiltonian(L, terms, Nup, Ndw)
            for fmt in ("G", "S", "N"):
                H = qcm.Hamiltonian(lab, sec, fmt)
                x = np.random.default_rng(1).normal(size=H.dim)
                assert H.dim == Hd.shape[0], (L, sec, fmt)
                y = H.apply(x)
                assert np.abs(y - Hd @ x).max() < 1e-12, (L, sec, fmt, np.abs(y - Hd @ x).max())
                del H
        qcm.erase_model_instance(lab)


def test_batched_solver_on_the_smallest_sectors():
    """the same one-state sectors solved by the batched solver (ground states
    and continued fractions for many instances at once)"""
    qcm.set_global_parameterC("GF_method", "F")
    U = 3.0
    labs = []
    for k in range(3):
        _label[0] += 1
        qcm.new_model_instanceC(_chain(1), {"U": U + k, "mu": (U + k) / 2}, "R0:N1:S1/R0:N1:S-1", _label[0])
        labs.append(_label[0])
    _label[0] += 1
    qcm.new_model_instanceC(_chain(3), {"t": 1.0, "U": 5.0, "mu": -2.0}, "R0:N0:S0", _label[0])
    labs.append(_label[0])
    qcm.solve_batch(labs)
    for k, lab in enumerate(labs[:3]):
        Uk = U + k
        assert abs(qcm.ground_state_solve(lab)[0] + Uk / 2) < 1e-12
        G = qcm.Green_function_grid(Z, label=lab)[:, 0, 0]
        ref = 0.5 / (Z + Uk / 2) + 0.5 / (Z - Uk / 2)
        assert np.abs(G - ref).max() < 1e-12, (k, np.abs(G - ref).max())
    lab = labs[3]
    assert abs(qcm.ground_state_solve(lab)[0]) < 1e-14
    G = qcm.Green_function_grid(Z, label=lab)
    h = _h(3, 1.0, -2.0)
    for iz, z in enumerate(Z):
        assert np.abs(G[iz] - np.linalg.inv(z * np.eye(3) - h)).max() < 1e-10
    for lab in labs:
        qcm.erase_model_instance(lab)
    qcm.set_global_parameterC("GF_method", "L")


def test_host_start_vector_solve():
    """the reference's host start vector (host_random_start) and the device
    one give the same ground state and Green function"""
    vals = {"t": 1.0, "U": 4.0, "mu": 2.0}
    out = []
    for host in (False, True):
        qcm.set_global_parameterC("host_random_start", host)
        lab = _instance(4, vals, "R0:N4:S0", "L")
        E0, _ = qcm.ground_state_solve(lab)
        out.append((E0, qcm.Green_function_grid(Z, label=lab)))
        qcm.erase_model_instance(lab)
    qcm.set_global_parameterC("host_random_start", False)
    assert abs(out[0][0] - out[1][0]) < 1e-10 * abs(out[0][0])
    assert round(out[1][0], 7) == -9.9531453  # examples/1D_4/GS.tsv:2
    assert np.abs(out[0][1] - out[1][1]).max() < 1e-8
