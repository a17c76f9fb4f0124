"""Batched multi-frequency synthesis on the GPU: every channel x stokes slice
of a context goes through one launch per step (uv grids, degridding, the
adjoint GEMM over a table of visibility ranges of all slices, one epilogue).
Checked against the FP64 oracle on shrunken versions of the multi-frequency
workload (uneven per-channel counts from flagging, spectral indices), for
the chain rule to I_nu_0 and to alpha, and with many slices."""
import numpy as np
import pytest

import oracle
from gpuvmem_amd.device import Context
from gpuvmem_amd.problem import make_mfs_problem

pytestmark = pytest.mark.gpu

CHI2_TOL = 1e-6
GRAD_TOL = 1e-5


def _small_mfs(n_spw=2, n_chan=4, npix=128, n_antennas=10, hours=0.2, seed=300):
    return make_mfs_problem(n_spw=n_spw, n_chan=n_chan, seed=seed, npix=npix,
                            n_antennas=n_antennas, ha_start_hours=-hours, ha_end_hours=hours,
                            n_gaussians=6, n_points=2, position_radius_frac=0.25,
                            sigma_min_pix=1.5, sigma_max_pix=4.0)


def _image(p, seed):
    rng = np.random.default_rng(seed)
    img = p.initial_image()
    img[0] += (rng.random((p.N, p.M)) * 40.0).astype(np.float32)
    img[1] = rng.uniform(-1.0, 3.0, (p.N, p.M)).astype(np.float32)
    return img


def _check(p, orc, mode, img, flag_opt, normalize=False):
    ctx = Context(p, adjoint=mode)
    d = ctx.image(img)
    v = ctx.chi2(d, flag_opt=flag_opt, normalize=normalize)
    ref, vr, clipped = orc.chi2(img, flag_opt=flag_opt, normalize=normalize)
    assert abs(v - ref) <= CHI2_TOL * abs(ref), (mode, flag_opt, v, ref)
    res = ctx.zeros(2 * p.M * p.N)
    ctx.dchi2(d, res, flag_opt=flag_opt, normalize=normalize)
    ctx.sync()
    g = res.download(np.float32, (2, p.N, p.M))
    ctx.close()
    gref = orc.dchi2(clipped, vr, flag_opt=flag_opt, normalize=normalize)
    k = flag_opt % 2
    err = np.linalg.norm(g[k] - gref[k]) / np.linalg.norm(gref[k])
    assert np.all(g[1 - k] == 0)
    return err


@pytest.mark.parametrize("mode", ["f16x3", "bf16x3", "f32", "direct"])
def test_mfs_batched_parity(mode):
    p = _small_mfs()
    counts = [u.shape[0] for _, u, _, _ in p.slices]
    print(f"case: MFS {len(counts)} channels {p.M}^2 counts {counts} mode {mode}", flush=True)
    assert len(set(counts)) == len(counts)  # uneven after flagging
    orc = oracle.Oracle(p)
    img = _image(p, 1)
    for flag in (0, 1):
        err = _check(p, orc, mode, img, flag)
        print(f"flag {flag}: gradient rel L2 {err:.3e}")
        assert err < GRAD_TOL, (mode, flag, err)


def test_mfs_many_slices_and_normalize():
    """256 channels (the workload's channel count) at a small size, with the
    per-slice 1/nvis normalisation."""
    p = _small_mfs(n_spw=4, n_chan=64, npix=64, n_antennas=5, hours=0.1)
    print(f"case: MFS 256 channels {p.M}^2 nvis {p.nvis_total}", flush=True)
    assert len(p.slices) == 256
    orc = oracle.Oracle(p)
    img = _image(p, 2)
    for flag in (0, 1):
        err = _check(p, orc, "f16x3", img, flag, normalize=(flag == 1))
        print(f"flag {flag}: gradient rel L2 {err:.3e}")
        assert err < GRAD_TOL
