"""The surrogate network's device kernels (bf16 MFMA GEMMs, fp32 accumulation)
against the FP64 oracle (oracle/mlp_ref.py).

Bars: with the oracle rounding where the device rounds (mimic), outputs and
gradients agree to fp32-accumulation level (outputs 1e-3 of their scale,
gradients 5e-3 relative per tensor, the loss 1e-4); against the unrounded
FP64 network, to bfloat16 level (3e-2 / 6e-2).  One Adam step moves every
parameter with a clear gradient by -lr sign(g), as the oracle's step does.
Training lowers the loss of a smooth regression by more than 10x."""
import numpy as np
import pytest

from oracle import mlp_ref as M
from pyqcm_amd import _lib
from pyqcm_amd.surrogate import MLP

pytestmark = pytest.mark.gpu

SHAPES = [(12, 641, 256, 4, 96), (5, 7, 64, 2, 33)]


@pytest.fixture(scope="module", autouse=True)
def _device():
    _lib.require_device()
    yield


def _net(n_in, n_out, width, depth, seed=11):
    net = MLP(n_in, n_out, width, depth, seed=seed)
    P = M.unflatten(net.get_weights(), n_in, n_out, width, depth)
    return net, P


@pytest.mark.parametrize("n_in,n_out,width,depth,B", SHAPES)
def test_forward_matches_oracle(n_in, n_out, width, depth, B):
    net, P = _net(n_in, n_out, width, depth)
    x = np.random.default_rng(5).normal(size=(B, n_in))
    y = net.predict(x)
    y_mim, _ = M.forward(P, x, mimic=True)
    y_ref, _ = M.forward(P, x, mimic=False)
    scale = np.max(np.abs(y_ref))
    assert np.max(np.abs(y - y_mim)) <= 1e-3 * scale, np.max(np.abs(y - y_mim)) / scale
    assert np.max(np.abs(y - y_ref)) <= 3e-2 * scale, np.max(np.abs(y - y_ref)) / scale


@pytest.mark.parametrize("n_in,n_out,width,depth,B", SHAPES)
def test_gradients_match_oracle(n_in, n_out, width, depth, B):
    net, P = _net(n_in, n_out, width, depth)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(B, n_in))
    t = rng.normal(size=(B, n_out))
    w = rng.uniform(0.5, 2.0, n_out)
    w[0] = 32.0
    net.set_scaling(loss_weight=w)
    loss, g = net.loss_grad(x, t)
    loss_m, G_m = M.loss_grad(P, x, M.f32(t), M.f32(w), mimic=True)
    loss_r, G_r = M.loss_grad(P, x, t, w, mimic=False)
    assert abs(loss - loss_m) <= 1e-4 * loss_m
    assert abs(loss - loss_r) <= 2e-2 * loss_r
    gl = M.unflatten(g, n_in, n_out, width, depth)
    for l, ((dW, db), (mW, mb), (rW, rb)) in enumerate(zip(gl, G_m, G_r)):
        for name, a, m, r in (("W", dW, mW, rW), ("b", db, mb, rb)):
            em = np.linalg.norm(a - m) / np.linalg.norm(m)
            er = np.linalg.norm(a - r) / np.linalg.norm(r)
            assert em <= 5e-3, (l, name, em)
            assert er <= 6e-2, (l, name, er)


def test_adam_step_matches_oracle():
    n_in, n_out, width, depth, B = 5, 7, 64, 2, 64
    net, P = _net(n_in, n_out, width, depth, seed=3)
    rng = np.random.default_rng(8)
    x = rng.normal(size=(B, n_in))
    t = rng.normal(size=(B, n_out))
    net.set_optimizer(lr=1e-3)
    _, g = net.loss_grad(x, t)
    hist, _ = net.train(x, t, epochs=1, batch=B, seed=0)   # one full batch: one step
    P1 = M.unflatten(net.get_weights(), n_in, n_out, width, depth)
    _, G = M.loss_grad(P, x, M.f32(t), np.ones(n_out), mimic=True)
    S = M.adam_state(P)
    R1 = M.adam_step(P, G, S, lr=1e-3)
    gl = M.unflatten(g, n_in, n_out, width, depth)
    for (W, b), (W1, b1), (Wr, br), (gW, gb) in zip(P, P1, R1, gl):
        for p, p1, pr, gg in ((W, W1, Wr, gW), (b, b1, br, gb)):
            clear = np.abs(gg) > 1e-5
            assert np.max(np.abs((p1 - pr)[clear])) <= 2e-6, np.max(np.abs((p1 - pr)[clear]))
            assert np.allclose((p1 - p)[clear], -1e-3 * np.sign(gg[clear]), rtol=2e-2, atol=1e-6)
    assert hist.shape == (1,) and np.isfinite(hist[0])


def test_training_lowers_the_loss():
    rng = np.random.default_rng(9)
    n, n_in, n_out = 4096, 12, 16
    x = rng.uniform(-1, 1, size=(n, n_in))
    A = rng.normal(size=(n_in, n_out))
    y = np.sin(x @ A) + 0.5 * (x[:, :1] ** 2)
    net = MLP(n_in, n_out, 128, 3, seed=1)
    net.set_scaling(x_mean=np.zeros(n_in), x_std=np.ones(n_in) * 0.6, y_mean=y.mean(0), y_std=y.std(0))
    net.set_optimizer(lr=2e-3)
    hist, ms = net.train(x, y, epochs=60, batch=512, seed=2)
    assert hist[-1] < hist[0] / 10, hist[[0, -1]]
    assert ms > 0
    # predictions in physical units follow the targets
    err = np.sqrt(np.mean((net.predict(x[:256]) - y[:256]) ** 2)) / np.std(y)
    assert err < 0.3, err


def test_bad_arguments_are_errors():
    with pytest.raises(_lib.EDError):
        MLP(3, 4, 12, 2)          # width not a multiple of 8
    net = MLP(3, 4, 16, 1)
    with pytest.raises(_lib.EDError):
        net.set_scaling(x_std=np.array([1.0, 0.0, 1.0]))
    with pytest.raises(ValueError):
        net.set_weights(np.zeros(3))
