"""The warm-start network's CPU oracle (oracle/nn_oracle.py) pinned by a
library routine (torch's float64 CPU convolutions, pooling, upsampling and
autograd) and by finite differences of the loss."""
import numpy as np
import pytest

import nn_oracle as nn

torch = pytest.importorskip("torch")
F = torch.nn.functional


def _t(a, grad=False):
    return torch.tensor(np.asarray(a, np.float64), requires_grad=grad)


@pytest.mark.parametrize("K,relu", [(3, True), (3, False), (1, False), (1, True)])
def test_conv_forward_and_backward_against_torch(K, relu):
    rng = np.random.default_rng(K + 2 * relu)
    x = rng.standard_normal((2, 3, 8, 6))
    w = rng.standard_normal((4, 3, K, K))
    b = rng.standard_normal(4)
    out = nn.conv_fwd(x, w, b, relu)
    tx, tw, tb = _t(x, True), _t(w, True), _t(b, True)
    to = F.conv2d(tx, tw, tb, padding=K // 2)
    if relu:
        to = torch.relu(to)
    assert np.allclose(out, to.detach().numpy(), rtol=1e-12, atol=1e-12)
    dout = rng.standard_normal(out.shape)
    to.backward(_t(dout))
    dx, dw, db = nn.conv_bwd(x, w, out, dout, relu)
    assert np.allclose(dx, tx.grad.numpy(), rtol=1e-12, atol=1e-12)
    assert np.allclose(dw, tw.grad.numpy(), rtol=1e-12, atol=1e-12)
    assert np.allclose(db, tb.grad.numpy(), rtol=1e-12, atol=1e-12)


def test_pool_and_upsample_against_torch():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 3, 8, 4))
    tx = _t(x, True)
    tp = F.avg_pool2d(tx, 2)
    assert np.allclose(nn.pool_fwd(x), tp.detach().numpy(), rtol=1e-14)
    g = rng.standard_normal(tp.shape)
    tp.backward(_t(g))
    assert np.allclose(nn.pool_bwd(g), tx.grad.numpy(), rtol=1e-14)
    tx2 = _t(x, True)
    tu = F.interpolate(tx2, scale_factor=2, mode="nearest")
    assert np.allclose(nn.up_fwd(x), tu.detach().numpy())
    g = rng.standard_normal(tu.shape)
    tu.backward(_t(g))
    assert np.allclose(nn.up_bwd(g), tx2.grad.numpy(), rtol=1e-13)


def _random_params(levels, base, seed):
    rng = np.random.default_rng(seed)
    return {name: (rng.standard_normal((co, ci, k, k)) * np.sqrt(2.0 / (ci * k * k)),
                   rng.standard_normal(co) * 0.1)
            for name, co, ci, k in nn.unet_shapes(levels, base)}


def _torch_unet(params, x, levels):
    tp = {k: (_t(w, True), _t(b, True)) for k, (w, b) in params.items()}
    h = _t(x)
    skips = []
    for l in range(levels):
        h = torch.relu(F.conv2d(h, *tp[f"e{l}a"], padding=1))
        h = torch.relu(F.conv2d(h, *tp[f"e{l}b"], padding=1))
        skips.append(h)
        if l < levels - 1:
            h = F.avg_pool2d(h, 2)
    d = skips[-1]
    for l in range(levels - 2, -1, -1):
        cat = torch.cat([F.interpolate(d, scale_factor=2, mode="nearest"), skips[l]], dim=1)
        d = torch.relu(F.conv2d(cat, *tp[f"d{l}a"], padding=1))
        d = torch.relu(F.conv2d(d, *tp[f"d{l}b"], padding=1))
    y = F.conv2d(d, *tp["out"])
    return y, tp


def test_unet_forward_and_backward_against_torch_autograd():
    levels, base = 3, 3
    params = _random_params(levels, base, 5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((2, 1, 16, 16))
    t = rng.standard_normal((2, 1, 16, 16))
    y, cache = nn.unet_forward(params, x, levels)
    ty, tp = _torch_unet(params, x, levels)
    assert np.allclose(y, ty.detach().numpy(), rtol=1e-11, atol=1e-12)
    loss, dy = nn.mse(y, t)
    tl = torch.mean((ty - _t(t)) ** 2)
    assert np.isclose(loss, tl.item(), rtol=1e-12)
    tl.backward()
    grads = nn.unet_backward(params, cache, dy, levels)
    for name in params:
        assert np.allclose(grads[name][0], tp[name][0].grad.numpy(), rtol=1e-9, atol=1e-12), name
        assert np.allclose(grads[name][1], tp[name][1].grad.numpy(), rtol=1e-9, atol=1e-12), name


def test_unet_gradient_finite_differences():
    levels, base = 2, 2
    params = _random_params(levels, base, 7)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((1, 1, 8, 8))
    t = rng.standard_normal((1, 1, 8, 8))

    def loss_of(p):
        y, _ = nn.unet_forward(p, x, levels)
        return nn.mse(y, t)[0]

    y, cache = nn.unet_forward(params, x, levels)
    _, dy = nn.mse(y, t)
    grads = nn.unet_backward(params, cache, dy, levels)
    for name in ("e0a", "d0a", "out"):
        w, b = params[name]
        idx = (0, 0, w.shape[2] // 2, w.shape[3] // 2)
        h = 1e-6
        wp = w.copy(); wp[idx] += h
        wm = w.copy(); wm[idx] -= h
        fd = (loss_of({**params, name: (wp, b)}) - loss_of({**params, name: (wm, b)})) / (2 * h)
        assert np.isclose(grads[name][0][idx], fd, rtol=1e-5, atol=1e-9), name
