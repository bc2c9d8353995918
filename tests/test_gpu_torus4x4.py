"""The large single solve: the 4x4 periodic Hubbard cluster at half filling
(t = 1, U = 8, mu = 4, sector N = 16, S = 0, dim 165 636 900, 1.3 GB per
vector) on the device.
  * Lanczos ground state (two-vector recurrence, eigenvector pass): E0 against
    the oracle's Lanczos on the same sector (tests/golden/torus4x4_u8_gs.json,
    written by tests/golden/make_torus4x4_gs.py), 1e-10 relative, and the
    Ritz vector's residual through the device product (itself checked
    against the oracle on this sector in test_gpu_parity.py);
  * the same with full FP64 reorthogonalisation against the Krylov basis
    kept in HBM;
  * Davidson on the 4 lowest states: ascending eigenvalues, the lowest equal
    to E0, orthonormal vectors with small residuals."""
import json
import os

import numpy as np
import pytest

from pyqcm_amd import _lib, qcm, workloads as W

pytestmark = pytest.mark.gpu

_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "torus4x4_u8_gs.json")
_SEC = "R0:N16:S0"
_label = [11000]


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    yield
    qcm.set_global_parameterC("lanczos_reorth", False)
    qcm.set_global_parameterC("lanczos_basis_budget_GB", -1.0)
    qcm.set_global_parameterC("accur_lanczos", 1e-12)


def _golden():
    with open(_GOLDEN) as f:
        return json.load(f)


def _instance():
    _label[0] += 1
    lab = _label[0]
    W.setup(W.torus(4, 4, U=8.0), label=lab, sectors=_SEC, model_name="torus4x4_%d" % lab)
    return lab


def _residual(H, E, v):
    v = v / np.linalg.norm(v)
    r = H.apply(v)
    r -= E * v
    return float(np.linalg.norm(r))


def test_torus4x4_lanczos_ground_state():
    g = _golden()
    lab = _instance()
    E0, sec = qcm.ground_state_solve(lab)
    print("E0", repr(E0), "golden", repr(g["E0"]), flush=True)
    assert abs(E0 - g["E0"]) <= 1e-10 * abs(g["E0"])
    qcm.erase_model_instance(lab)


def test_torus4x4_lanczos_eigenvector():
    """the Ritz vector from the Krylov basis kept in HBM (default budget: the
    available device memory; the sector needs ~170 vectors, 225 GB): E0 and
    the residual through the device product"""
    g = _golden()
    lab = _instance()
    H = qcm.Hamiltonian(lab, _SEC, "G")
    start = np.random.default_rng(3).normal(size=H.dim)
    E0, gs, nit = H.lanczos_gs(start)
    del start
    r = _residual(H, E0, gs)
    print("E0", repr(E0), nit, "iterations, residual", r, flush=True)
    assert abs(E0 - g["E0"]) <= 1e-10 * abs(g["E0"])
    assert r < 1e-7
    del gs
    qcm.erase_model_instance(lab)


def test_torus4x4_lanczos_reorthogonalised():
    """full reorthogonalisation, Krylov basis held in HBM (up to 150 vectors,
    200 GB); stop rule 1e-9 on the Ritz residual keeps the basis inside it"""
    g = _golden()
    qcm.set_global_parameterC("lanczos_reorth", True)
    qcm.set_global_parameterC("lanczos_basis_budget_GB", 200.0)
    qcm.set_global_parameterC("accur_lanczos", 1e-9)
    try:
        lab = _instance()
        E0, _ = qcm.ground_state_solve(lab)
        print("E0 (reorth)", repr(E0), flush=True)
        assert abs(E0 - g["E0"]) <= 1e-10 * abs(g["E0"])
        qcm.erase_model_instance(lab)
    finally:
        qcm.set_global_parameterC("lanczos_reorth", False)
        qcm.set_global_parameterC("lanczos_basis_budget_GB", -1.0)
        qcm.set_global_parameterC("accur_lanczos", 1e-12)


def test_torus4x4_davidson_four_lowest():
    g = _golden()
    lab = _instance()
    H = qcm.Hamiltonian(lab, _SEC, "G")
    ev, vec, nit = H.davidson(4)
    print("Davidson", nit, "iterations:", ev, flush=True)
    assert np.all(np.diff(ev) >= -1e-12)
    assert abs(ev[0] - g["E0"]) <= 1e-9 * abs(g["E0"])
    gram = vec @ vec.T
    assert np.abs(gram - np.eye(4)).max() < 1e-8
    for k in range(4):
        r = _residual(H, ev[k], vec[k])
        print("residual", k, r, flush=True)
        assert r < 1e-4
    del vec
    qcm.erase_model_instance(lab)
