"""The learned surrogate in the VCA sweep: the network trained on the ED's own
(parameters -> E0, continued fractions) pairs proposes the stationary point of
each problem; the reference's Newton-Raphson procedure then runs from there
with every evaluation a full ED solve, so each accepted point is the exact
solver's.  The same procedure from the default start is the plain path.

Bars: the two paths find the same stationary points (|dOmega| < 1e-8, |dM| <
1e-3) and the surrogate path needs fewer full solves (the count of both is
printed; the trained network and its corpus are described by the weights
file's metadata)."""
import os

import numpy as np
import pytest

from pyqcm_amd import _lib, qcm, vca_surrogate as V

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pyqcm_amd", "data",
                       "plaq8b_surrogate.npz")


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    qcm.set_global_parameterC("Hamiltonian_format", "G")
    qcm.set_global_parameterC("GF_method", "L")
    yield


def test_surrogate_predicts_held_out_solves():
    """on 16 parameter points outside the corpus (another seed): E0 within 2 %
    of the exact solver's everywhere, and mean absolute errors of E0 and of the
    fractions' first level within twice those on the 2000 corpus rows held out
    of training (recorded with the weights): the network does as well on new
    points as on its own held-out rows"""
    import ast
    net, meta = V.load(WEIGHTS)
    meta = ast.literal_eval(meta)
    X = V.sample(16, seed=1234)
    Y = net.predict(X)
    eE, e0 = [], []
    for k in range(len(X)):
        lab = V.new_instance(X[k], 7601)
        y, _ = V.exact_targets(lab)
        assert abs(Y[k, 0] - y[0]) < 2e-2 * abs(y[0]), (k, Y[k, 0], y[0])
        eE.append(abs(Y[k, 0] - y[0]))
        _, a_p, b_p = V.unpack(Y[k])
        _, a_x, b_x = V.unpack(y)
        e0.append(0.5 * (np.abs(a_p[:, 0] - a_x[:, 0]).mean() + np.abs(b_p[:, 0] - b_x[:, 0]).mean()))
    print("E0 mae %.4f (held out %.4f), level-0 mae %.4f (held out %.4f)" %
          (np.mean(eE), meta["holdout_E0_mae"], np.mean(e0), meta["holdout_cf_mae_level0"]), flush=True)
    assert np.mean(eE) < 2.0 * meta["holdout_E0_mae"]
    assert np.mean(e0) < 2.0 * meta["holdout_cf_mae_level0"]


def test_surrogate_vca_finds_the_same_stationary_points_with_fewer_solves():
    net, meta = V.load(WEIGHTS)
    problems = V.sweep_problems(4)
    grid = V.OmegaGrid(nk=24, w=(2.0, 20.0, 24, 24, 24))
    r = V.vca_compare(net, problems, grid)
    print("plain solves", r["solves_plain"], "surrogate solves", r["solves_surrogate"],
          "ratio %.2f" % (r["solves_plain"] / max(1, r["solves_surrogate"])), flush=True)
    for k, (p, s) in enumerate(zip(r["plain"], r["surrogate"])):
        print(k, "M plain %.6f (%d it)" % (p["x"][0], p["iterations"]),
              "M surrogate %.6f (%d it, start %.4f)" % (s["x"][0], s["iterations"], r["starts"][k][0]),
              "dOmega %.2e" % abs(p["omega"] - s["omega"]), flush=True)
        assert abs(p["omega"] - s["omega"]) < 1e-8
        assert abs(p["x"][0] - s["x"][0]) < 1e-3
    assert r["solves_surrogate"] < r["solves_plain"]
