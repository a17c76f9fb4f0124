"""The two batchings of the image-to-grid transform (wide blocks for many
slices, narrow blocks for one small image) run the same arithmetic: chi2,
residuals and gradients are bit-identical, for one slice and for several."""
import numpy as np
import pytest

from gpuvmem_amd.device import Context
from gpuvmem_amd.problem import small_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("npix,nslices", [(64, 1), (128, 1), (256, 1), (512, 1), (128, 3)])
def test_wide_and_narrow_fft_batching_agree_bitwise(npix, nslices):
    p = small_problem(npix=npix, n_antennas=8, seed=51)
    if nslices > 1:
        f0, u, v, w = p.all_slices()[0]
        p.slices = [(f0, u, v, w)] * nslices
    print(f"case: fft batching {npix}^2 x{nslices} nvis={p.nvis}", flush=True)
    img = p.initial_image()
    img[0] += (np.random.default_rng(2).random((p.N, p.M)) * 10).astype(np.float32)
    out = {}
    for mode in ("wide", "narrow"):
        ctx = Context(p)
        ctx.set_fft_mode(mode)
        d = ctx.image(img)
        v = ctx.chi2(d)
        r = ctx.zeros(2 * p.M * p.N)
        ctx.dchi2(d, r)
        ctx.sync()
        out[mode] = (v, ctx.residuals(), r.download(np.float32))
        ctx.close()
    (vw, rw, gw), (vn, rn, gn) = out["wide"], out["narrow"]
    assert vw == vn
    assert np.array_equal(rw, rn)
    assert np.array_equal(gw, gn)
