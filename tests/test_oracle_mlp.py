"""The FP64 network oracle (oracle/mlp_ref.py) pinned by the mathematics:
central finite differences of the loss, closed forms of GELU, the bfloat16
rounding rule on known values, and the first Adam step."""
import math

import numpy as np

from oracle import mlp_ref as M


def _params(rng, n_in=3, n_out=4, width=8, depth=2):
    return [(rng.normal(size=(o, i)) / math.sqrt(i), 0.1 * rng.normal(size=o))
            for (o, i) in M.layer_shapes(n_in, n_out, width, depth)]


def test_gelu_closed_forms():
    assert M.gelu(np.array(0.0)) == 0.0
    assert abs(M.gelu_grad(np.array(0.0)) - 0.5) < 1e-15
    # Phi(1) = 0.5 (1 + erf(1 / sqrt 2)) = 0.8413447460685429
    assert abs(M.gelu(np.array(1.0)) - 0.8413447460685429) < 1e-15
    assert abs(M.gelu(np.array(-1.0)) - (-1.0 + 0.8413447460685429)) < 1e-15
    assert abs(M.gelu(np.array(12.0)) - 12.0) < 1e-12 and abs(M.gelu(np.array(-12.0))) < 1e-12
    z = np.linspace(-4, 4, 81)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_grad(z))) < 1e-9


def test_bf16_rounding_rule():
    assert M.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0                 # tie -> even (down)
    assert M.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6  # tie -> even (up)
    assert M.bf16(np.array([1.0 + 2.0 ** -8 + 2.0 ** -12]))[0] == 1.0 + 2.0 ** -7
    assert M.bf16(np.array([np.pi]))[0] == 3.140625
    assert M.bf16(np.array([-np.pi]))[0] == -3.140625
    x = np.random.default_rng(0).normal(size=1000)
    r = M.bf16(x)
    assert np.max(np.abs(r - x) / np.abs(x)) <= 2.0 ** -8


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    P = _params(rng)
    x = rng.normal(size=(5, 3))
    t = rng.normal(size=(5, 4))
    w = np.array([4.0, 1.0, 0.5, 2.0])
    loss, grads = M.loss_grad(P, x, t, w)
    flat = M.flatten(P)
    g_flat = M.flatten(grads)
    h = 1e-6
    fd = np.zeros_like(flat)
    for k in range(flat.size):
        e = np.zeros_like(flat)
        e[k] = h
        lp, _ = M.loss_grad(M.unflatten(flat + e, 3, 4, 8, 2), x, t, w)
        lm, _ = M.loss_grad(M.unflatten(flat - e, 3, 4, 8, 2), x, t, w)
        fd[k] = (lp - lm) / (2 * h)
    assert np.max(np.abs(fd - g_flat)) < 1e-7 * max(1.0, np.max(np.abs(g_flat)))
    # the loss itself: mean weighted squared error
    y, _ = M.forward(P, x)
    assert abs(loss - np.sum(w * (y - t) ** 2) / 5) < 1e-14


def test_mimic_rounds_only_where_stated():
    rng = np.random.default_rng(2)
    P = [(M.bf16(W), M.f32(b)) for (W, b) in _params(rng)]
    x = M.bf16(rng.normal(size=(6, 3)))
    y0, _ = M.forward(P, x)
    y1, _ = M.forward(P, x, mimic=True)
    # only the hidden activations are rounded in between: a bf16-sized change
    assert 0 < np.max(np.abs(y1 - y0)) < 2e-2 * np.max(np.abs(y0))


def test_first_adam_step_moves_by_lr_times_sign():
    rng = np.random.default_rng(3)
    P = _params(rng)
    x = rng.normal(size=(7, 3))
    t = rng.normal(size=(7, 4))
    _, G = M.loss_grad(P, x, t, np.ones(4))
    S = M.adam_state(P)
    P1 = M.adam_step(P, G, S, lr=1e-3)
    for (W, b), (W1, b1), (gW, gb) in zip(P, P1, G):
        for p, p1, g in ((W, W1, gW), (b, b1, gb)):
            big = np.abs(g) > 1e-4
            assert np.allclose((p1 - p)[big], -1e-3 * np.sign(g[big]), rtol=1e-3, atol=0)
    assert S["t"] == 1
