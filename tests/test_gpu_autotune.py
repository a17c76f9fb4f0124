"""The adjoint's range length (K tile) and its autotuner: every candidate
range length keeps the gradient within the bar (small size against the
FP64 oracle everywhere, full size against FP64 direct sums on sampled
pixels at the shortest and longest default candidates); the autotuner
picks one of its candidates, caches it per shape in the process and in the
GVM_TUNE_CACHE file."""
import numpy as np
import pytest

import oracle
from gpuvmem_amd.device import Context
from gpuvmem_amd.problem import make_problem, small_problem

pytestmark = pytest.mark.gpu


def _img(p, seed=0):
    rng = np.random.default_rng(seed)
    img = p.initial_image()
    img[0] += (rng.random((p.N, p.M)) * 20.0).astype(np.float32)
    return img


@pytest.mark.parametrize("mode", ["bf16x3", "f16x3", "f32"])
def test_every_range_length_keeps_parity(mode):
    p = small_problem(npix=128, n_antennas=12, seed=31)
    print(f"case: range lengths {mode} {p.M}^2 nvis={p.nvis}", flush=True)
    orc = oracle.Oracle(p)
    img = _img(p)
    ref, vr, clipped = orc.chi2(img)
    gref = orc.dchi2(clipped, vr)[0]
    ctx = Context(p, adjoint=mode)
    d = ctx.image(img)
    ctx.chi2(d)
    for sv in (256, 1536, 3072, 8192, 16384):
        ctx.set_split_vis(sv)
        r = ctx.zeros(2 * p.M * p.N)
        ctx.dchi2(d, r)
        ctx.sync()
        g = r.download(np.float32, (2, p.N, p.M))[0]
        err = np.linalg.norm(g - gref) / np.linalg.norm(gref)
        print(f"split_vis {sv}: gradient rel L2 {err:.2e}")
        assert err < 1e-5
    ctx.close()


def test_autotuner_picks_and_caches(tmp_path, monkeypatch):
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("GVM_TUNE_CACHE", str(cache))
    p = small_problem(npix=128, n_antennas=11, seed=32)
    print(f"case: autotune {p.M}^2 nvis={p.nvis}", flush=True)
    cands = [512, 1024, 2048]
    ctx = Context(p)
    d = ctx.image(_img(p))
    ctx.chi2(d)
    r = ctx.autotune(cands)
    print(r)
    assert r["split_vis"] in cands and not r["cached"]
    assert set(r["ms"]) == set(cands) and all(v > 0 for v in r["ms"].values())
    assert ctx.split_vis() == r["split_vis"]
    # the tuned context still evaluates; a second context of the same shape hits the cache
    res = ctx.zeros(2 * p.M * p.N)
    ctx.dchi2(d, res)
    ctx.sync()
    ctx.close()
    ctx2 = Context(p)
    ctx2.chi2(ctx2.image(_img(p)))
    r2 = ctx2.autotune(cands)
    assert r2["cached"] and r2["split_vis"] == r["split_vis"]
    ctx2.close()
    text = cache.read_text()
    assert f"= {r['split_vis']}" in text and f"M={p.M}" in text


@pytest.fixture(scope="module")
def full():
    return make_problem(1)


@pytest.mark.parametrize("mode", ["bf16x3", "f16x3"])
def test_full_size_gradient_at_range_lengths(full, mode):
    """The shortest and longest default candidates, the previous default
    (3072) and the default (6144)."""
    p = full
    print(f"case: full size range lengths {mode} {p.M}^2 nvis={p.nvis}", flush=True)
    orc = oracle.Oracle(p)
    img = p.initial_image()
    img[0] = np.maximum(p.sky / p.fg_scale * 0.5, p.minpix).astype(np.float32)
    _, vr, _ = orc.chi2(img)
    rng = np.random.default_rng(5)
    inside = np.argwhere(p.noise < p.noise_cut)
    pix = inside[rng.choice(len(inside), 32, replace=False)]
    gref = orc.dchi2_pixels(vr, pix)
    ctx = Context(p, adjoint=mode)
    d = ctx.image(img)
    ctx.chi2(d)
    for sv in (1536, 3072, 6144, 8192):
        ctx.set_split_vis(sv)
        r = ctx.zeros(2 * p.M * p.N)
        ctx.dchi2(d, r)
        ctx.sync()
        g = r.download(np.float32, (2, p.N, p.M))[0]
        gs = np.array([g[i, j] for i, j in pix])
        err = np.linalg.norm(gs - gref) / np.linalg.norm(gref)
        print(f"split_vis {sv}: sampled gradient rel L2 {err:.2e}")
        assert err < 1e-5
    ctx.close()
