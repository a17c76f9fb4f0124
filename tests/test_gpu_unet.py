"""The warm-start network's device kernels and the whole U-Net (forward,
backward, Adam) against the FP64 CPU oracle of the same layers."""
import ctypes as C

import numpy as np
import pytest

import nn_oracle as nn
from gpuvmem_amd import native
from gpuvmem_amd.device import Context, DeviceBuffer
from gpuvmem_amd.problem import small_problem
from gpuvmem_amd.warmstart import UNet

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    p = small_problem(npix=64, n_antennas=5, seed=41)
    c = Context(p, add_data=False)
    yield c
    c.close()


def _dev(a):
    return DeviceBuffer.from_array(np.ascontiguousarray(a, np.float32))


def _get(buf, shape):
    return buf.download(np.float32)[: int(np.prod(shape))].reshape(shape).astype(np.float64)


@pytest.mark.parametrize("K,relu", [(3, True), (3, False), (1, False)])
def test_conv_kernels_match_oracle(ctx, K, relu):
    print(f"case: conv K={K} relu={relu}", flush=True)
    rng = np.random.default_rng(K + 10 * relu)
    B, Ci, Co, H, W = 2, 5, 4, 12, 8
    x = rng.standard_normal((B, Ci, H, W)).astype(np.float32)
    w = (rng.standard_normal((Co, Ci, K, K)) * 0.3).astype(np.float32)
    b = rng.standard_normal(Co).astype(np.float32)
    dx, dw, db = _dev(x), _dev(w), _dev(b)
    out = ctx.zeros(B * Co * H * W)
    native.check(ctx.lib.gvm_nn_conv_fwd(ctx.ctx, dx.ptr, dw.ptr, db.ptr, out.ptr, B, Ci, Co, H, W,
                                         K, int(relu)), "fwd")
    ref = nn.conv_fwd(x.astype(np.float64), w.astype(np.float64), b.astype(np.float64), relu)
    ctx.sync()
    o = _get(out, ref.shape)
    assert np.allclose(o, ref, rtol=1e-5, atol=1e-5)
    g = rng.standard_normal(ref.shape).astype(np.float32)
    dg = _dev(g)
    din, dwg, dbg = ctx.zeros(x.size), ctx.zeros(w.size), ctx.zeros(b.size)
    native.check(ctx.lib.gvm_nn_conv_bwd(ctx.ctx, dx.ptr, dw.ptr, out.ptr if relu else None, dg.ptr,
                                         din.ptr, dwg.ptr, dbg.ptr, B, Ci, Co, H, W, K, int(relu)),
                 "bwd")
    rx, rw, rb = nn.conv_bwd(x.astype(np.float64), w.astype(np.float64), o, g.astype(np.float64),
                             relu)
    ctx.sync()
    assert np.allclose(_get(din, x.shape), rx, rtol=1e-5, atol=1e-5)
    assert np.allclose(_get(dwg, w.shape), rw, rtol=1e-5, atol=1e-4)
    assert np.allclose(_get(dbg, b.shape), rb, rtol=1e-5, atol=1e-5)


def _bf16_round(w):
    u = np.ascontiguousarray(w, np.float32).view(np.uint32).astype(np.uint64)
    h = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return h, (h.astype(np.uint32) << 16).view(np.float32)


def test_bf16_weight_conv_matches_oracle(ctx):
    print("case: conv bf16 weights", flush=True)
    rng = np.random.default_rng(9)
    B, Ci, Co, H, W, K = 2, 6, 5, 10, 8, 3
    x = rng.standard_normal((B, Ci, H, W)).astype(np.float32)
    w = (rng.standard_normal((Co, Ci, K, K)) * 0.3).astype(np.float32)
    b = rng.standard_normal(Co).astype(np.float32)
    wh, wr = _bf16_round(w)
    # the device rounding equals the host's
    dwf = _dev(w)
    dwh = DeviceBuffer(2 * w.size)
    native.check(ctx.lib.gvm_nn_to_bf16(ctx.ctx, dwf.ptr, dwh.ptr, w.size), "to_bf16")
    ctx.sync()
    assert np.array_equal(dwh.download(np.uint16), wh.ravel())
    dx, db = _dev(x), _dev(b)
    out = ctx.zeros(B * Co * H * W)
    native.check(ctx.lib.gvm_nn_conv_fwd_bf16(ctx.ctx, dx.ptr, dwh.ptr, db.ptr, out.ptr, B, Ci, Co,
                                              H, W, K, 1), "fwd")
    ref = nn.conv_fwd(x.astype(np.float64), wr.astype(np.float64), b.astype(np.float64), True)
    ctx.sync()
    o = _get(out, ref.shape)
    assert np.allclose(o, ref, rtol=1e-5, atol=1e-5)
    g = rng.standard_normal(ref.shape).astype(np.float32)
    din, dwg, dbg = ctx.zeros(x.size), ctx.zeros(w.size), ctx.zeros(b.size)
    native.check(ctx.lib.gvm_nn_conv_bwd_bf16(ctx.ctx, dx.ptr, dwh.ptr, out.ptr, _dev(g).ptr,
                                              din.ptr, dwg.ptr, dbg.ptr, B, Ci, Co, H, W, K, 1),
                 "bwd")
    rx, rw, rb = nn.conv_bwd(x.astype(np.float64), wr.astype(np.float64), o, g.astype(np.float64),
                             True)
    ctx.sync()
    assert np.allclose(_get(din, x.shape), rx, rtol=1e-5, atol=1e-5)
    assert np.allclose(_get(dwg, w.shape), rw, rtol=1e-5, atol=1e-4)
    assert np.allclose(_get(dbg, b.shape), rb, rtol=1e-5, atol=1e-5)


def test_pool_up_copy_mse_adam_match_oracle(ctx):
    print("case: pool/up/copy/mse/adam", flush=True)
    rng = np.random.default_rng(2)
    B, Cc, H, W = 2, 3, 8, 6
    x = rng.standard_normal((B, Cc, H, W)).astype(np.float32)
    dx = _dev(x)
    po = ctx.zeros(x.size // 4)
    native.check(ctx.lib.gvm_nn_pool_fwd(ctx.ctx, dx.ptr, po.ptr, B, Cc, H, W), "pool")
    ctx.sync()
    assert np.allclose(_get(po, (B, Cc, H // 2, W // 2)), nn.pool_fwd(x), rtol=1e-6, atol=1e-6)
    g = rng.standard_normal((B, Cc, H // 2, W // 2)).astype(np.float32)
    pb = ctx.zeros(x.size)
    native.check(ctx.lib.gvm_nn_pool_bwd(ctx.ctx, _dev(g).ptr, pb.ptr, B, Cc, H, W), "pool bwd")
    ctx.sync()
    assert np.allclose(_get(pb, x.shape), nn.pool_bwd(g), rtol=1e-6, atol=1e-7)
    # upsample into channels [2, 5) of a 7-channel tensor, and back
    Cd = 7
    up = ctx.zeros(B * Cd * 4 * H * W)
    native.check(ctx.lib.gvm_nn_up_fwd(ctx.ctx, dx.ptr, up.ptr, B, Cc, Cd, 2, H, W), "up")
    ctx.sync()
    u = _get(up, (B, Cd, 2 * H, 2 * W))
    assert np.array_equal(u[:, 2:5], nn.up_fwd(x)) and np.all(u[:, :2] == 0) and np.all(u[:, 5:] == 0)
    gu = rng.standard_normal((B, Cd, 2 * H, 2 * W)).astype(np.float32)
    ub = ctx.zeros(x.size)
    native.check(ctx.lib.gvm_nn_up_bwd(ctx.ctx, _dev(gu).ptr, ub.ptr, B, Cc, Cd, 2, H, W), "up bwd")
    ctx.sync()
    assert np.allclose(_get(ub, x.shape), nn.up_bwd(gu[:, 2:5].astype(np.float64)), rtol=1e-6,
                       atol=1e-6)
    # channel copy
    cp = ctx.zeros(B * 6 * H * W)
    native.check(ctx.lib.gvm_nn_copy_channels(ctx.ctx, dx.ptr, cp.ptr, B, 2, Cc, 1, 6, 3, H * W),
                 "copy")
    ctx.sync()
    c6 = _get(cp, (B, 6, H, W))
    assert np.array_equal(c6[:, 3:5], x[:, 1:3]) and np.all(c6[:, :3] == 0)
    # mse
    t = rng.standard_normal(x.shape).astype(np.float32)
    dy = ctx.zeros(x.size)
    native.check(ctx.lib.gvm_nn_mse(ctx.ctx, dx.ptr, _dev(t).ptr, dy.ptr, x.size, 3), "mse")
    out = (C.c_double * 4)()
    native.check(ctx.lib.gvm_values_read(ctx.ctx, 4, out), "values")
    loss, dref = nn.mse(x.astype(np.float64), t.astype(np.float64))
    assert np.isclose(out[3], loss, rtol=1e-6)
    assert np.allclose(_get(dy, x.shape), dref, rtol=1e-5)
    # adam: two steps against the textbook update
    p = rng.standard_normal(50).astype(np.float32)
    gr = rng.standard_normal(50).astype(np.float32)
    dp, dgr, dm, dv = _dev(p), _dev(gr), ctx.zeros(50), ctx.zeros(50)
    m = np.zeros(50); v = np.zeros(50); pr = p.astype(np.float64)
    for step in (1, 2):
        native.check(ctx.lib.gvm_nn_adam(ctx.ctx, dp.ptr, dgr.ptr, dm.ptr, dv.ptr, 50, 1e-2, 0.9,
                                         0.999, 1e-8, step), "adam")
        m = 0.9 * m + 0.1 * gr; v = 0.999 * v + 0.001 * gr.astype(np.float64) ** 2
        pr -= 1e-2 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    ctx.sync()
    assert np.allclose(_get(dp, (50,)), pr, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("bf16", [False, True])
def test_unet_forward_backward_match_oracle(ctx, bf16):
    levels, base, B, H = 3, 4, 2, 16
    print(f"case: unet levels={levels} base={base} {B}x{H}^2 bf16={bf16}", flush=True)
    net = UNet(ctx, levels=levels, base=base, batch=B, H=H, W=H, seed=3, bf16=bf16)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((B, 1, H, H)).astype(np.float32)
    t = rng.standard_normal((B, 1, H, H)).astype(np.float32)
    net.set_input(x)
    net.forward()
    loss = net.loss(_dev(t))
    net.backward()
    y = net.output().astype(np.float64)
    params = net.params_dict()
    yref, cache = nn.unet_forward(params, x.astype(np.float64), levels)
    lref, dy = nn.mse(yref, t.astype(np.float64))
    err_y = np.abs(y - yref).max() / np.abs(yref).max()
    print(f"output max rel {err_y:.2e}; loss {loss:.6e} vs {lref:.6e}")
    assert err_y < 1e-5 and abs(loss - lref) <= 1e-5 * lref
    gref = nn.unet_backward(params, cache, dy, levels)
    grads = net.grads_dict()
    for name in params:
        for k in (0, 1):
            a, r = grads[name][k], gref[name][k]
            err = np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30)
            assert err < 1e-4, (name, k, err)
    # a second backward of the same batch is bit-identical (fixed-order reductions)
    g1 = net.grads.download(np.float32)
    net.forward(); net.loss(_dev(t)); net.backward()
    assert np.array_equal(net.grads.download(np.float32), g1)
    # inference on fewer images than the batch gives the same outputs
    y_all = net.output()
    net.set_input(x[1:2])
    net.forward(1)
    assert np.array_equal(net.output(1)[0], y_all[1])
    net.close()
