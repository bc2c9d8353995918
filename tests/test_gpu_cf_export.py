"""Continued fractions out of a solved instance and into another one: the
exported fractions (a_n, b_n as evaluated) of both spin blocks, set on a fresh
instance of the same model, give the same Green function; the exported
fractions are the oracle's LanczosGreen for the same orbital pairs."""
import numpy as np
import pytest

from pyqcm_amd import _lib, qcm, workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    qcm.set_global_parameterC("GF_method", "F")
    yield
    qcm.set_global_parameterC("GF_method", "L")


def test_export_import_round_trip(oracle):
    wl = W.hubbard_chain(4, U=4.0, mu=2.0)
    name = W.setup(wl, label=9701, model_name="cfx_chain4")
    E0, sec = qcm.ground_state_solve(9701)
    z = W.matsubara(64, 20.0)
    G = [qcm.Green_function_grid(z, s, 9701) for s in (0, 1)]
    qcm.new_model_instanceC(name, dict(wl.values), wl.sector, 9702)
    for s in (0, 1):
        a, b, ln = qcm.continued_fractions(9701, s)
        assert a.shape[0] == 4 * 5 and ln.min() > 0
        qcm.set_continued_fractions(9702, E0, ":".join(sec.split(":")[:3]), s, a, b, ln)
    for s in (0, 1):
        assert np.abs(qcm.Green_function_grid(z, s, 9702) - G[s]).max() < 1e-12
    # the diagonal pair (o1 = o2 = 0), electron part, against the oracle
    a, b, ln = qcm.continued_fractions(9701, 0)
    M = oracle.Model(wl.L, wl.terms())
    x = np.random.default_rng(3).normal(size=36)
    E0r, gs, _, _ = M.lanczos_gs(2, 2, x)
    gs = gs / np.linalg.norm(gs)
    psi = oracle.create_or_destroy(wl.L, 1, 0, 2, 2, gs)
    nrm = psi @ psi
    ar, br = M.lanczos_green(3, 2, psi / np.sqrt(nrm))
    ar, br = oracle.cf_transform(ar, br, E0r, nrm, True)
    m = min(ln[0], len(ar), 12)
    assert np.abs(a[0, :m] - ar[:m]).max() < 1e-9
    assert np.abs(b[0, :m] - br[:m]).max() < 1e-9
    qcm.erase_model_instance(9701)
    qcm.erase_model_instance(9702)
