"""Pins of the network oracle (oracle/cnn_oracle.py), CPU only: the
convolution against scipy's correlate2d, the back-propagated gradient
against central finite differences, Adam against its one-step closed form."""
import numpy as np
import pytest
from scipy.signal import correlate2d

from oracle import cnn_oracle as C


def test_conv_matches_scipy_correlate2d():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((3, 11, 7))
    w = rng.standard_normal((3, 3, 3, 5))
    b = rng.standard_normal(5)
    z = C.conv(a, w, b)
    for co in range(5):
        want = b[co] + sum(correlate2d(a[ci], w[ci, :, :, co], mode="same", boundary="fill") for ci in range(3))
        np.testing.assert_allclose(z[co], want, rtol=1e-12, atol=1e-12)


def test_forward_applies_relu_between_layers_only():
    rng = np.random.default_rng(1)
    ch = [2, 4, 1]
    p = rng.standard_normal(C.num_params(ch))
    x = rng.standard_normal((2, 6, 5))
    acts = C.forward(p, ch, x)
    assert np.all(acts[1] >= 0)
    assert np.any(acts[2] < 0)
    (w1, b1), (w2, b2) = C.unpack(p, ch)
    np.testing.assert_allclose(acts[2], C.conv(np.maximum(C.conv(x, w1, b1), 0), w2, b2))


@pytest.mark.parametrize("masked", [False, True])
def test_backprop_matches_finite_differences(masked):
    rng = np.random.default_rng(2)
    ch = [2, 4, 4, 1]
    p = rng.standard_normal(C.num_params(ch)) * 0.5
    x = rng.standard_normal((2, 7, 9))
    t = rng.standard_normal((7, 9))
    m = (rng.random((7, 9)) > 0.3).astype(float) if masked else None
    f, g = C.loss_grad(p, ch, x, t, m)
    assert f == pytest.approx(C.loss(p, ch, x, t, m), rel=1e-14)
    h = 1e-6
    for k in range(p.size):
        e = np.zeros_like(p)
        e[k] = h
        fd = (C.loss(p + e, ch, x, t, m) - C.loss(p - e, ch, x, t, m)) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9), k


def test_adam_first_step_closed_form():
    g = np.array([3.0, -0.5, 1e-3])
    p0 = np.array([1.0, 2.0, 3.0])
    p, m, v = C.adam_step(p0, g, np.zeros(3), np.zeros(3), 1, lr=0.1, eps=0.0)
    np.testing.assert_allclose(p, p0 - 0.1 * np.sign(g))
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)
