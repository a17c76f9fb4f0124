"""The benchmark's larger workloads at their own sizes, against the FP64
oracle:

* 2048^2 with 4,354,560 visibilities (baselines to 1e6 wavelengths): chi2,
  and the gradient of every adjoint form on sampled pixels against FP64
  direct sums (the precision comparison of the split forms);
* 512^2 I_nu_0 + alpha over 4 x 64 channels with uneven flagging (61.4 M
  visibilities in one context): chi2 for both chain-rule branches, and the
  gradient on sampled pixels against FP64 direct sums over every channel;
* the warm-start network at its size (6 levels, 32 base channels, 512^2
  input, BF16 weights): one forward and backward pass against the FP64
  network oracle.
"""
import numpy as np
import pytest

import nn_oracle as nn
import oracle
from gpuvmem_amd.device import Context, DeviceBuffer
from gpuvmem_amd.problem import make_mfs_problem, make_problem
from gpuvmem_amd.warmstart import UNet

pytestmark = pytest.mark.gpu

CHI2_TOL = 1e-6
GRAD_TOL = 1e-5


def _sky_image(p):
    img = p.initial_image()
    img[0] = np.maximum(p.sky / p.fg_scale * 0.5, p.minpix).astype(np.float32)
    return img


def _sample(p, n, seed):
    rng = np.random.default_rng(seed)
    inside = np.argwhere(p.noise < p.noise_cut)
    return inside[rng.choice(len(inside), n, replace=False)]


def test_precision_forms_at_2048sq_4354560_visibilities():
    p = make_problem(2)
    print(f"case: {p.M}^2 nvis={p.nvis}", flush=True)
    assert (p.M, p.nvis) == (2048, 4354560)
    uvmax = np.hypot(p.uvw[:, 0], p.uvw[:, 1]).max()
    print(f"max |uv| {uvmax:.4e} wavelengths", flush=True)
    assert 0.9e6 < uvmax <= 1.0e6 * (1 + 1e-6)
    img = _sky_image(p)
    orc = oracle.Oracle(p)
    ref, vr, clipped = orc.chi2(img)
    pix = _sample(p, 32, 20)
    gref = orc.dchi2_pixels(vr, pix)
    errs = {}
    for mode in ("bf16x3", "f16x3", "f32", "direct"):
        print(f"case: {p.M}^2 adjoint {mode}", flush=True)
        ctx = Context(p, adjoint=mode)
        d = ctx.image(img)
        v = ctx.chi2(d)
        assert abs(v - ref) <= CHI2_TOL * abs(ref), (mode, v, ref)
        res = ctx.zeros(2 * p.M * p.N)
        ctx.dchi2(d, res)
        ctx.sync()
        g = res.download(np.float32, (2, p.N, p.M))[0]
        ctx.close()
        gs = np.array([g[i, j] for i, j in pix], np.float64)
        errs[mode] = np.linalg.norm(gs - gref) / np.linalg.norm(gref)
        print(f"{mode}: chi2 rel {abs(v - ref) / abs(ref):.2e}; sampled gradient rel L2 "
              f"{errs[mode]:.3e}", flush=True)
    assert all(e < GRAD_TOL for e in errs.values()), errs


@pytest.fixture(scope="module")
def mfs_full():
    p = make_mfs_problem()
    return p


def test_multi_frequency_workload_61m_visibilities(mfs_full):
    p = mfs_full
    counts = [s[1].shape[0] for s in p.slices]
    print(f"case: MFS {p.M}^2 {len(counts)} channels nvis={p.nvis_total}", flush=True)
    assert len(counts) == 256 and len(set(counts)) > 200
    assert 55e6 < p.nvis_total < 65e6
    rng = np.random.default_rng(30)
    img = _sky_image(p)
    img[1] = rng.uniform(-1.0, 3.0, (p.N, p.M)).astype(np.float32)
    orc = oracle.Oracle(p)
    pix = _sample(p, 12, 31)
    ctx = Context(p)
    d = ctx.image(img)
    for flag in (0, 1):
        d.upload(img)
        v = ctx.chi2(d, flag_opt=flag)
        ref, vr, clipped = orc.chi2(img, flag_opt=flag)
        print(f"flag {flag}: chi2 {v:.10e} vs {ref:.10e} (rel {abs(v - ref) / abs(ref):.2e})",
              flush=True)
        assert abs(v - ref) <= CHI2_TOL * abs(ref)
        np.testing.assert_array_equal(d.download(np.float32, img.shape),
                                      clipped.astype(np.float32))
        res = ctx.zeros(2 * p.M * p.N)
        ctx.dchi2(d, res, flag_opt=flag)
        ctx.sync()
        g = res.download(np.float32, (2, p.N, p.M))
        res.free()
        gref = orc.dchi2_pixels_all(clipped, vr, pix, flag_opt=flag)
        k = flag % 2
        gs = np.array([g[k, i, j] for i, j in pix], np.float64)
        err = np.linalg.norm(gs - gref[:, k]) / np.linalg.norm(gref[:, k])
        print(f"flag {flag}: sampled gradient rel L2 {err:.3e}", flush=True)
        assert err < GRAD_TOL
        assert np.all(g[1 - k] == 0)
    ctx.close()


def test_warm_start_network_at_size():
    """6 levels, 32 base channels (32..1024 channels), a 512^2 input, BF16
    weights with FP32 accumulation: outputs and every parameter gradient
    against the FP64 oracle evaluated with the same (BF16-rounded) weights."""
    levels, base, H = 6, 32, 512
    print(f"case: unet levels={levels} base={base} 1x{H}^2 bf16", flush=True)
    p = make_problem(0, npix=64)
    ctx = Context(p, add_data=False)
    net = UNet(ctx, levels=levels, base=base, batch=1, H=H, W=H, seed=40, bf16=True)
    rng = np.random.default_rng(41)
    x = rng.standard_normal((1, 1, H, H)).astype(np.float32)
    t = rng.standard_normal((1, 1, H, H)).astype(np.float32)
    net.set_input(x)
    net.forward()
    loss = net.loss(DeviceBuffer.from_array(t))
    net.backward()
    y = net.output().astype(np.float64)
    params = net.params_dict()
    grads = net.grads_dict()
    net.close()
    ctx.close()
    yref, cache = nn.unet_forward(params, x.astype(np.float64), levels)
    lref, dy = nn.mse(yref, t.astype(np.float64))
    err_y = np.abs(y - yref).max() / np.abs(yref).max()
    print(f"output max rel {err_y:.2e}; loss {loss:.8e} vs {lref:.8e}", flush=True)
    assert err_y < 1e-5 and abs(loss - lref) <= 1e-5 * lref
    gref = nn.unet_backward(params, cache, dy, levels)
    worst = 0.0
    for name in params:
        for k in (0, 1):
            a, r = grads[name][k], gref[name][k]
            err = np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30)
            worst = max(worst, err)
            assert err < 1e-4, (name, k, err)
    print(f"parameter gradients: worst rel L2 {worst:.2e}", flush=True)
