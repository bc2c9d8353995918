"""Concurrent work on one device.
  * Calls on different instance labels from several host threads (the
    reference's OpenMP loop over clusters, src_qcm/lattice_model_instance.cpp:107-113,
    :160) run at once, each thread queuing on a stream of its own: they give
    the results of the same calls made one after the other.
  * The band-Lanczos Green function's hole and electron parts run on two
    host threads and streams (gf_streams 2): the same Lehmann representation
    as one after the other (gf_streams 1).
  * A label used from several threads in turn sees the work of the previous
    call complete (every call returns with its device work done)."""
import threading

import numpy as np
import pytest

from pyqcm_amd import _lib, qcm
from pyqcm_amd import workloads as W

pytestmark = pytest.mark.gpu

_label = [9900]
_models = set()
_Z = W.matsubara(16, 50.0)


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    qcm.set_global_parameterC("GF_method", "L")
    yield
    qcm.set_global_parameterC("gf_streams", 2)


def _new(wl, values=None):
    if wl.name not in _models:
        qcm.new_model(wl.name, wl.n_sites, wl.n_bath)
        for op, (typ, el) in wl.operators.items():
            qcm.new_operator(wl.name, op, typ, [(r + 1, c + 1, v) for (r, c, v) in el])
        _models.add(wl.name)
    _label[0] += 1
    vals = dict(wl.values)
    vals.update(values or {})
    qcm.new_model_instanceC(wl.name, vals, wl.sector, _label[0])
    return _label[0]


def _solve(lab):
    E0, sec = qcm.ground_state_solve(lab)
    G = qcm.Green_function_grid(_Z, label=lab)
    ave = qcm.cluster_averages(lab)
    return E0, sec, G, ave


def _cases():
    # sectors of 4 900 to 63 504 states: a 2x2 plaquette with 4 bath
    # orbitals, a 10-site chain, and the 2x2 plaquette with 6 bath orbitals
    return [(W.plaquette_bath(seed=3, n_bath=4), {"U": 4.0}),
            (W.plaquette_bath(seed=3, n_bath=4), {"U": 6.0, "mu": 3.0}),
            (W.hubbard_chain(10, U=4.0, mu=2.0), {}),
            (W.plaquette_bath(seed=5, n_bath=6), {"U": 8.0})]


def _same(a, b):
    assert a[1] == b[1]
    assert abs(a[0] - b[0]) <= 1e-11 * abs(b[0]), (a[0], b[0])
    assert np.abs(a[2] - b[2]).max() < 1e-9, np.abs(a[2] - b[2]).max()
    for k in b[3]:
        assert abs(a[3][k][0] - b[3][k][0]) < 1e-9, (k, a[3][k], b[3][k])


def test_concurrent_labels_match_serial_calls():
    serial = []
    for wl, v in _cases():
        lab = _new(wl, v)
        serial.append(_solve(lab))
        qcm.erase_model_instance(lab)
    labs = [_new(wl, v) for wl, v in _cases()]
    got, errs = {}, []

    def run(i, lab):
        try:
            got[i] = _solve(lab)
        except Exception as e:  # reported below
            errs.append((i, repr(e)))

    ths = [threading.Thread(target=run, args=(i, lab)) for i, lab in enumerate(labs)]
    for t in ths:
        t.start()
    for t in ths:
        t.join(300)
    assert not errs, errs
    assert all(not t.is_alive() for t in ths) and len(got) == len(labs)
    for i in range(len(labs)):
        _same(got[i], serial[i])
    for lab in labs:
        qcm.erase_model_instance(lab)


@pytest.mark.parametrize("n_bath", [4, 8])
def test_gf_streams_give_the_same_lehmann_representation(n_bath):
    wl = W.plaquette_bath(n_bath=n_bath)
    out = {}
    for k in (1, 2):
        qcm.set_global_parameterC("gf_streams", k)
        lab = _new(wl)
        out[k] = _solve(lab) + qcm.qmatrix(lab)
        qcm.erase_model_instance(lab)
    qcm.set_global_parameterC("gf_streams", 2)
    _same(out[2], out[1])
    e1, Q1 = out[1][4], out[1][5]
    e2, Q2 = out[2][4], out[2][5]
    assert e1.shape == e2.shape
    assert np.abs(e1 - e2).max() < 1e-10
    assert np.abs(Q1 @ Q1.T - Q2 @ Q2.T).max() < 1e-10


def test_one_label_from_several_threads_in_turn():
    """a solve on one thread, the evaluation on another, the averages on a
    third: each call finds the previous one's device work complete"""
    wl = W.plaquette_bath(seed=3, n_bath=4)
    ref = _new(wl)
    want = _solve(ref)
    qcm.erase_model_instance(ref)
    lab = _new(wl)
    res = {}

    def step(name, fn):
        t = threading.Thread(target=lambda: res.__setitem__(name, fn()))
        t.start()
        t.join(120)
        assert not t.is_alive()

    step("gs", lambda: qcm.ground_state_solve(lab))
    step("gf", lambda: qcm.Green_function_grid(_Z, label=lab))
    step("ave", lambda: qcm.cluster_averages(lab))
    _same((res["gs"][0], res["gs"][1], res["gf"], res["ave"]), want)
    qcm.erase_model_instance(lab)


def test_many_threads_mixed_calls():
    """eight threads, each creating, solving, evaluating and erasing its own
    instances in a loop (with threads ending and new ones starting, so
    streams, partial-sum arenas and staging buffers are handed on), while
    the main thread defines a new model: every result equals the serial one"""
    cases = _cases()
    serial = []
    for wl, v in cases:
        lab = _new(wl, v)
        serial.append(_solve(lab))
        qcm.erase_model_instance(lab)
    lock = threading.Lock()
    errs = []

    def work(t):
        try:
            for r in range(3):
                i = (t + r) % len(cases)
                wl, v = cases[i]
                with lock:
                    lab = _new(wl, v)
                got = _solve(lab)
                qcm.erase_model_instance(lab)
                _same(got, serial[i])
        except Exception as e:  # reported below
            errs.append((t, repr(e)))

    for wave in range(2):
        ths = [threading.Thread(target=work, args=(t,)) for t in range(8)]
        for t in ths:
            t.start()
        if wave == 0:
            qcm.new_model("conc_extra", 2, 0)
            qcm.new_operator("conc_extra", "t", "one-body", [(1, 2, -1.0), (3, 4, -1.0)])
        for t in ths:
            t.join(600)
        assert all(not t.is_alive() for t in ths)
    assert not errs, errs[:3]
