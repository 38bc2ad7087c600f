"""The network's HIP kernels against the float64 oracle (oracle/cnn_oracle.py):
forward, loss, parameter gradient and Adam; and training that learns."""
import numpy as np
import pytest

from oracle import cnn_oracle as C
from seispie_amd import cnn
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _case(ch, H, W, seed):
    rng = np.random.default_rng(seed)
    p = cnn.init_params(ch, seed)
    p[-1] = 0.3  # a nonzero output bias
    x = rng.standard_normal((ch[0], H, W)).astype(np.float32)
    t = rng.standard_normal((H, W)).astype(np.float32)
    m = (rng.random((H, W)) > 0.25).astype(np.float32)
    return p, x, t, m


@pytest.mark.parametrize("ch,H,W", [([2, 8, 8, 1], 37, 45), ([2, 32, 32, 32, 1], 29, 70), ([3, 16, 4, 2], 9, 33)])
def test_forward_matches_oracle(ch, H, W):
    p, x, _, _ = _case(ch, H, W, 1)
    print("case: cnn forward", ch, H, W, flush=True)
    net = cnn.Net(ch, params=p)
    y = net.forward(x)
    want = C.forward(p.astype(np.float64), ch, x)[-1]
    assert y.shape == want.shape
    err = rel_l2(y, want)
    print("forward rel L2 %.2e" % err)
    assert err <= 1e-5
    np.testing.assert_array_equal(net.params(), p)
    net.close()


@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("ch,H,W", [([2, 8, 8, 1], 37, 45), ([2, 32, 32, 32, 1], 40, 64)])
def test_loss_and_gradient_match_oracle(ch, H, W, masked):
    p, x, t, m = _case(ch, H, W, 2)
    m = m if masked else None
    print("case: cnn gradient", ch, H, W, masked, flush=True)
    net = cnn.Net(ch, params=p)
    f = net.loss_grad(x, t, m)
    g = net.grad()
    fw, gw = C.loss_grad(p.astype(np.float64), ch, x, t, m)
    assert f == pytest.approx(fw, rel=1e-5)
    o = 0
    for ci, co in zip(ch[:-1], ch[1:]):
        nw = ci * 9 * co
        ew = rel_l2(g[o:o + nw], gw[o:o + nw])
        eb = rel_l2(g[o + nw:o + nw + co], gw[o + nw:o + nw + co])
        print("layer %d->%d: dW rel L2 %.2e, db %.2e" % (ci, co, ew, eb))
        assert ew <= 1e-5 and eb <= 1e-5
        o += nw + co
    net.close()


def test_adam_steps_match_oracle():
    ch = [2, 8, 8, 1]
    p, x, t, m = _case(ch, 21, 34, 3)
    net = cnn.Net(ch, params=p)
    pp = p.astype(np.float64)
    mm = np.zeros_like(pp)
    vv = np.zeros_like(pp)
    for step in range(1, 4):
        net.loss_grad(x, t, m)
        g = net.grad().astype(np.float64)
        net.adam_step(1e-3)
        pp, mm, vv = C.adam_step(pp, g, mm, vv, step, 1e-3)
        np.testing.assert_allclose(net.params(), pp, rtol=1e-5, atol=1e-7)
    net.close()


def test_training_learns_a_smoothing_map():
    """The network learns target = 3x3 box mean of channel 0 times
    (1 + channel 1): the loss drops by two orders of magnitude."""
    from scipy.ndimage import uniform_filter
    rng = np.random.default_rng(4)
    ch = [2, 16, 16, 1]
    x = rng.standard_normal((2, 48, 64)).astype(np.float32)
    x[1] = 0.5 * np.abs(x[1])
    t = (uniform_filter(x[0], 3, mode="constant") * (1 + x[1])).astype(np.float32)
    net = cnn.Net(ch, seed=5)
    first = net.loss_grad(x, t)
    for _ in range(600):
        last = net.loss_grad(x, t)
        net.adam_step(3e-3)
    print("loss %.3e -> %.3e" % (first, last))
    assert last < 1e-2 * first
    net.close()
