"""Batched solves (one workgroup per sector or per continued fraction) against
the per-instance path on the same parameters and against dense
diagonalization.  Bars: E0 1e-10 relative; G(z) 1e-8 on a Matsubara grid,
both spin blocks."""
import dataclasses

import numpy as np
import pytest

from pyqcm_amd import qcm, workloads as W, _lib
from tests.dense_ref import dense_hamiltonian

pytestmark = pytest.mark.gpu

_label = [5000]
_models = set()


def _new_label():
    _label[0] += 1
    return _label[0]


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    yield
    qcm.set_global_parameterC("GF_method", "L")


def _model(wl):
    name = "batch_" + wl.name
    if name not in _models:
        qcm.new_model(name, wl.n_sites, wl.n_bath)
        for op, (typ, el) in wl.operators.items():
            qcm.new_operator(name, op, typ, [(r + 1, c + 1, v) for (r, c, v) in el])
        _models.add(name)
    return name


def _dense_E0(wl, values):
    w = dataclasses.replace(wl, values=values)
    e0 = np.inf
    for sec in wl.sector.split("/"):
        N, S = qcm.sector_parse(sec)[:2]
        Nup, Ndw = (N + S) // 2, (N - S) // 2
        e0 = min(e0, np.linalg.eigvalsh(dense_hamiltonian(w.L, w.terms(), Nup, Ndw))[0])
    return e0


def _compare(wl, points, gf_method, z, dense=True):
    qcm.set_global_parameterC("GF_method", gf_method)
    name = _model(wl)
    single, batch = [], []
    for vals in points:
        la, lb = _new_label(), _new_label()
        qcm.new_model_instanceC(name, vals, wl.sector, la)
        qcm.new_model_instanceC(name, vals, wl.sector, lb)
        single.append(la)
        batch.append(lb)
    for la in single:
        qcm.ground_state_solve(la)
        qcm.Green_function_solveC(la)
    qcm.solve_batch(batch, green_function=True)
    for vals, la, lb in zip(points, single, batch):
        Ea, sa = qcm.ground_state_solve(la)
        Eb, sb = qcm.ground_state_solve(lb)
        assert abs(Ea - Eb) <= 1e-10 * abs(Ea), (vals, Ea, Eb)
        assert sa.split(":")[:3] == sb.split(":")[:3]
        if dense:
            Ed = _dense_E0(wl, vals)
            assert abs(Eb - Ed) <= 1e-10 * abs(Ed), (vals, Eb, Ed)
        for spin in (0, 1):
            Ga = qcm.Green_function_grid(z, spin_down=spin, label=la)
            Gb = qcm.Green_function_grid(z, spin_down=spin, label=lb)
            assert np.abs(Ga - Gb).max() < 1e-8, (vals, spin, np.abs(Ga - Gb).max())
    # all batched instances evaluated together equal their one-by-one evaluation
    for spin in (0, 1):
        Gs = qcm.Green_function_batch(z, batch, spin_down=spin)
        for lb, G in zip(batch, Gs):
            assert np.abs(G - qcm.Green_function_grid(z, spin_down=spin, label=lb)).max() < 1e-12
    for la, lb in zip(single, batch):
        qcm.erase_model_instance(la)
        qcm.erase_model_instance(lb)


def test_batch_2x2_neel_points_continued_fractions():
    _compare(W.vca_cluster(2, 2), W.vca_points(12, seed=4), "F", W.matsubara(32, 20.0))


def test_batch_2x3_neel_points_continued_fractions():
    _compare(W.vca_cluster(2, 3), W.vca_points(4, seed=5), "F", W.matsubara(32, 20.0))


def test_batch_ground_states_with_band_lanczos_green_functions():
    """GF_method 'L': batched ground states, band-Lanczos Green functions per instance."""
    _compare(W.vca_cluster(2, 2), W.vca_points(3, seed=6), "L", W.matsubara(32, 20.0))


def test_batch_large_sector_four_wave_kernel():
    """A 10-site chain at half filling (dim 63 504; N +- 1 sectors 52 920) runs
    in the four-wave workgroup."""
    wl = W.hubbard_chain(10, U=4.0, mu=2.0)
    _compare(wl, [{"t": 1.0, "U": 4.0, "mu": 2.0}, {"t": 1.0, "U": 6.0, "mu": 3.1}], "F", W.matsubara(16, 20.0),
             dense=False)


def test_batch_48400_state_sector_global_memory(oracle):
    """The 2x3 cluster with 6 bath orbitals at N = 6 (220 x 220 = 48 400
    states; the fractions run in the N +- 1 sectors of 14 520 and 108 900
    states): every item goes through the global-memory workgroups.  E0 against
    the oracle's Lanczos, E0 and G against the per-instance path."""
    wl = W.vca_cluster_bath(2, 3, 6)
    pts = [{**wl.values, **v} for v in W.vca_points(2, seed=4)]   # bath values kept
    _compare(wl, pts, "F", W.matsubara(16, 20.0), dense=False)
    for vals in pts:
        w = dataclasses.replace(wl, values=vals)
        M = oracle.Model(w.L, w.terms())
        x = np.random.default_rng(0).normal(size=M.dim(3, 3))
        E0r = M.lanczos_gs(3, 3, x)[0]
        lab = _new_label()
        qcm.new_model_instanceC(_model(wl), vals, wl.sector, lab)
        qcm.solve_batch([lab])
        E0 = qcm.ground_state_solve(lab)[0]
        assert abs(E0 - E0r) <= 1e-10 * abs(E0r), (vals, E0, E0r)
        qcm.erase_model_instance(lab)


def test_batch_mixed_formats_and_empty():
    """An empty batch is a no-op; instances built with the CSR format ('S')
    are solved one by one inside the same call, next to batched 'G' ones, and
    all agree with their own solves."""
    qcm.solve_batch([])
    qcm.set_global_parameterC("GF_method", "F")
    wl = W.vca_cluster(2, 2)
    name = _model(wl)
    pts = W.vca_points(4, seed=9)
    z = W.matsubara(16, 20.0)
    labs, refs = [], []
    for k, vals in enumerate(pts):
        qcm.set_global_parameterC("Hamiltonian_format", "S" if k % 2 else "G")
        la, lb = _new_label(), _new_label()
        qcm.new_model_instanceC(name, vals, wl.sector, la)
        qcm.new_model_instanceC(name, vals, wl.sector, lb)
        labs.append(la)
        refs.append(lb)
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    qcm.solve_batch(labs)
    for la, lb in zip(labs, refs):
        Ea, _ = qcm.ground_state_solve(la)
        Eb, _ = qcm.ground_state_solve(lb)
        assert abs(Ea - Eb) <= 1e-10 * abs(Eb)
        for spin in (0, 1):
            Ga = qcm.Green_function_grid(z, spin_down=spin, label=la)
            Gb = qcm.Green_function_grid(z, spin_down=spin, label=lb)
            assert np.abs(Ga - Gb).max() < 1e-8
        qcm.erase_model_instance(la)
        qcm.erase_model_instance(lb)
