"""FP64 restatement of the learned surrogate's network (test infrastructure only:
the product trains and evaluates it on the device; this module is the checker
of those kernels and is imported by tests/ alone).

The network: x -> depth hidden layers h = gelu(W h + b) of one width -> linear
output.  GELU is the exact erf form 0.5 z (1 + erf(z / sqrt 2)).  The loss is
sum_o w_o (y_o - t_o)^2 averaged over the batch (normalized units), and the
optimizer Adam (Kingma and Ba, 2015) with bias-corrected moments.

No reference counterpart exists (pyqcm has no learned model).  The functions
here are pinned by the mathematics: central finite differences of the loss
fix the gradients, the closed forms of GELU at special points and its
asymptotes fix the activation, and one Adam step from zero moments moves each
parameter by lr * sign(g) (up to eps).

``mimic=True`` rounds where the device rounds: inputs, weights, hidden
activations and the GEMM-side gradients to bfloat16 (round to nearest even),
so that the device and this module differ only by fp32 accumulation order.
"""
import numpy as np
from scipy.special import erf

SQRT1_2 = 0.70710678118654752
INV_SQRT_2PI = 0.39894228040143268


def bf16(a):
    """Round to bfloat16 (nearest, ties to even) through float32; returns float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64).reshape(np.shape(a))


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * SQRT1_2))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z * SQRT1_2)) + z * INV_SQRT_2PI * np.exp(-0.5 * z * z)


def layer_shapes(n_in, n_out, width, depth):
    shapes = []
    for l in range(depth + 1):
        fan_in = n_in if l == 0 else width
        fan_out = n_out if l == depth else width
        shapes.append((fan_out, fan_in))
    return shapes


def unflatten(w, n_in, n_out, width, depth):
    """[(W, b)] from the flat weight order of the C-ABI (per layer W row-major, then b)."""
    out, k = [], 0
    for (o, i) in layer_shapes(n_in, n_out, width, depth):
        W = np.asarray(w[k:k + o * i], dtype=np.float64).reshape(o, i)
        k += o * i
        b = np.asarray(w[k:k + o], dtype=np.float64)
        k += o
        out.append((W, b))
    assert k == len(w)
    return out


def flatten(params):
    return np.concatenate([np.concatenate([W.ravel(), b]) for (W, b) in params])


def forward(params, x, mimic=False):
    """Returns (y, cache); x normalized, shape (B, n_in)."""
    r = bf16 if mimic else (lambda a: a)
    h = r(x)
    zs, hs = [], [h]
    for l, (W, b) in enumerate(params):
        Wr = r(W)
        z = h @ Wr.T + (f32(b) if mimic else b)
        if l < len(params) - 1:
            zs.append(z)
            h = r(gelu(z))
            hs.append(h)
        else:
            y = z
    return y, (zs, hs)


def loss_grad(params, x, t, w_loss, mimic=False):
    """Mean weighted squared error and its gradient [(dW, db)] (same shapes)."""
    r = bf16 if mimic else (lambda a: a)
    B = x.shape[0]
    y, (zs, hs) = forward(params, x, mimic)
    res = y - t
    loss = float(np.sum(w_loss[None, :] * res * res) / B)
    dz = r(2.0 * w_loss[None, :] * res / B)
    grads = [None] * len(params)
    for l in range(len(params) - 1, -1, -1):
        W, b = params[l]
        grads[l] = (dz.T @ hs[l], dz.sum(axis=0))
        if l == 0:
            break
        da = dz @ r(W)
        dz = r(da * gelu_grad(zs[l - 1]))
    return loss, grads


def adam_step(params, grads, state, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update; state = dict(t, m, v) (lists of arrays), updated in place."""
    state["t"] += 1
    t = state["t"]
    c1, c2 = 1.0 - b1 ** t, 1.0 - b2 ** t
    new = []
    for k, ((W, b), (gW, gb)) in enumerate(zip(params, grads)):
        out = []
        for j, (p, g) in enumerate(((W, gW), (b, gb))):
            key = 2 * k + j
            m = b1 * state["m"][key] + (1 - b1) * g
            v = b2 * state["v"][key] + (1 - b2) * g * g
            state["m"][key], state["v"][key] = m, v
            out.append(p - lr * (m / c1) / (np.sqrt(v / c2) + eps))
        new.append(tuple(out))
    return new


def adam_state(params):
    z = []
    for (W, b) in params:
        z += [np.zeros_like(W), np.zeros_like(b)]
    return {"t": 0, "m": list(z), "v": [a.copy() for a in z]}
