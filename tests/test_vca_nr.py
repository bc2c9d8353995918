"""The Newton-Raphson procedure of the VCA drivers (the reference's
pyqcm/vca.py:241-430, batched over independent problems) on closed forms:
the quadratic fit is exact for quadratics, and the procedure converges to
the stationary point with 1 + 2n + n(n-1)/2 evaluations per iteration plus
the final one."""
import numpy as np

from pyqcm_amd import vca_surrogate as V


def test_fit_is_exact_for_quadratics():
    rng = np.random.default_rng(0)
    A = rng.normal(size=(2, 2))
    H = A @ A.T + np.eye(2)
    g0 = rng.normal(size=2)
    f = lambda x: 0.7 + g0 @ x + 0.5 * x @ H @ x
    x = np.array([0.3, -0.2])
    pts = V.nr_points(x, np.array([0.01, 0.02]))
    assert len(pts) == 1 + 4 + 1
    val, g, Hf = V.nr_fit(pts, [f(q) for q in pts])
    assert abs(val - f(x)) < 1e-10
    assert np.allclose(g, g0 + H @ x, atol=1e-8)
    assert np.allclose(Hf, H, atol=1e-6)


def test_procedure_finds_the_stationary_points_and_counts_evaluations():
    f1 = lambda x: 3.0 * (x[0] - 0.2) ** 2 + 0.5 * (x[0] - 0.2) ** 3 - 1.0
    f2 = lambda x: 1.5 * (x[0] - 0.05) ** 2 - 2.0 * (x[0] - 0.05) ** 4
    fs = [f1, f2]
    calls = []

    def evaluate(req):
        calls.append(len(req))
        return [fs[k](x) for (k, x) in req]

    out = V.newton_raphson(evaluate, [[0.1], [0.25]], step=0.01, accur=1e-4)
    # within the procedure's accuracy (accur = 1e-4 on the last move; the fit over
    # +-step of a cubic biases the gradient by O(step^2))
    assert abs(out[0]["x"][0] - 0.2) < 1e-5 and abs(out[1]["x"][0] - 0.05) < 1e-5
    for o in out:
        assert o["evaluations"] == 3 * o["iterations"] + 1
    assert abs(out[0]["omega"] + 1.0) < 1e-9
    # batched: every round evaluates all pending problems together
    assert sum(calls) == sum(o["evaluations"] for o in out)
    # a start at the solution converges in one iteration
    one = V.newton_raphson(evaluate, [[0.2]], step=0.01, accur=1e-4)[0]
    assert one["iterations"] == 1 and one["evaluations"] == 4
