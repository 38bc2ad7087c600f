"""CPU oracle of the step conditioner's network (test infrastructure only:
used by tests/ and never by the product path).

float64 NumPy restatement of include/seispie_cnn.h: 3x3 convolutions with
zero padding, ReLU after every layer but the last, parameters per layer as
W[cin][3][3][cout] then b[cout]; the masked mean-squared-error loss, its
parameter gradient by back-propagation, and one bias-corrected Adam step.
No reference counterpart (icui/seispie has no learned component).

Pinned in tests/test_oracle_pins.py: the convolution against
scipy.signal.correlate2d, the back-propagated gradient against central
finite differences of the loss, Adam against its closed form for one step.
"""
import numpy as np


def layer_shapes(channels):
    return [(channels[l], channels[l + 1]) for l in range(len(channels) - 1)]


def num_params(channels):
    return sum(ci * 9 * co + co for ci, co in layer_shapes(channels))


def unpack(params, channels):
    out, o = [], 0
    for ci, co in layer_shapes(channels):
        w = np.asarray(params[o:o + ci * 9 * co], np.float64).reshape(ci, 3, 3, co)
        o += ci * 9 * co
        b = np.asarray(params[o:o + co], np.float64)
        o += co
        out.append((w, b))
    return out


def init_params(channels, seed=0):
    """He-normal weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in layer_shapes(channels):
        parts.append(rng.standard_normal(ci * 9 * co) * np.sqrt(2.0 / (ci * 9)))
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)


def conv(a, w, b):
    """a [cin][H][W], w [cin][3][3][cout], b [cout] -> z [cout][H][W]."""
    ci, H, W = a.shape
    ap = np.zeros((ci, H + 2, W + 2))
    ap[:, 1:-1, 1:-1] = a
    z = np.broadcast_to(np.asarray(b, np.float64)[:, None, None], (w.shape[3], H, W)).copy()
    for dy in range(3):
        for dx in range(3):
            z += np.einsum("chw,co->ohw", ap[:, dy:dy + H, dx:dx + W], w[:, dy, dx, :])
    return z


def forward(params, channels, x):
    """Returns the activations a_0 = x, ..., a_L (a_L is the output z_L)."""
    acts = [np.asarray(x, np.float64)]
    layers = unpack(params, channels)
    for l, (w, b) in enumerate(layers):
        z = conv(acts[-1], w, b)
        acts.append(np.maximum(z, 0.0) if l < len(layers) - 1 else z)
    return acts


def loss(params, channels, x, t, m=None):
    y = forward(params, channels, x)[-1][0]
    m = np.ones_like(y) if m is None else np.asarray(m, np.float64)
    return float(np.sum(m * (y - t) ** 2) / np.sum(m))


def loss_grad(params, channels, x, t, m=None):
    """(loss, flat parameter gradient) by back-propagation."""
    acts = forward(params, channels, x)
    layers = unpack(params, channels)
    y = acts[-1][0]
    m = np.ones_like(y) if m is None else np.asarray(m, np.float64)
    ms = np.sum(m)
    L = len(layers)
    dz = (2.0 * m * (y - t) / ms)[None]
    grads = [None] * L
    for l in range(L - 1, -1, -1):
        a = acts[l]
        ci, H, W = a.shape
        ap = np.zeros((ci, H + 2, W + 2))
        ap[:, 1:-1, 1:-1] = a
        w, _ = layers[l]
        gw = np.zeros_like(w)
        for dy in range(3):
            for dx in range(3):
                gw[:, dy, dx, :] = np.einsum("chw,ohw->co", ap[:, dy:dy + H, dx:dx + W], dz)
        grads[l] = (gw, dz.sum(axis=(1, 2)))
        if l > 0:
            # dL/da_{l} = sum_t W[:, t, :] applied to dz shifted back by d_t
            dzp = np.zeros((dz.shape[0], H + 2, W + 2))
            dzp[:, 1:-1, 1:-1] = dz
            da = np.zeros((ci, H, W))
            for dy in range(3):
                for dx in range(3):
                    da += np.einsum("ohw,co->chw", dzp[:, 2 - dy:2 - dy + H, 2 - dx:2 - dx + W], w[:, dy, dx, :])
            dz = da * (a > 0)
    flat = np.concatenate([np.concatenate([g.reshape(-1), gb]) for g, gb in grads])
    return float(np.sum(m * (y - t) ** 2) / ms), flat


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam update; step counts from 1. Returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
