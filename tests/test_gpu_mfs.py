"""Batched multi-frequency synthesis on the GPU: every channel x stokes slice
of a context goes through one launch per step (uv grids, degridding, the
adjoint GEMM over a table of visibility ranges of all slices, one epilogue).
Checked against the FP64 oracle on shrunken versions of the multi-frequency
workload (uneven per-channel counts from flagging, spectral indices), for
the chain rule to I_nu_0 and to alpha, and with many slices."""
import ctypes as C

import numpy as np
import pytest

import oracle
from gpuvmem_amd import native
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


class _PriorTerm(C.Structure):
    """gvm_prior_term of include/gvm_hip.h."""
    _fields_ = [("kind", C.c_int), ("index", C.c_int), ("image_to_add", C.c_int),
                ("active", C.c_int), ("prior", C.c_float), ("eta", C.c_float),
                ("lambda_", C.c_float), ("pad", C.c_float)]


@pytest.mark.parametrize("n_spw,flag_opt,normalize", [(1, 0, False), (1, 1, True), (2, 2, False)])
def test_value_launch_is_bitwise_the_separate_launches(n_spw, flag_opt, normalize):
    """gvm_values_enqueue (chi2 and three priors; for up to four slices the
    degrid and the prior partials share one launch and one launch finishes
    them, for more slices each takes its own) leaves the same values bit for
    bit as gvm_chi2_enqueue followed by the one-term prior launches, for 3
    slices and for 6; chi2 also matches the FP64 oracle."""
    p = _small_mfs(n_spw=n_spw, n_chan=3, npix=64, n_antennas=8)
    nsl = len(p.slices)
    print(f"case: value launch {nsl} slices {p.M}^2 flag_opt {flag_opt} normalize {normalize}",
          flush=True)
    img = _image(p, 7)
    terms = (_PriorTerm * 3)()
    for t, (kind, index, prior, eta) in enumerate([(0, 0, 0.001, -1.0), (1, 1, 0.0, 0.0),
                                                    (2, 0, 0.0, 0.0)]):
        terms[t].kind, terms[t].index, terms[t].image_to_add = kind, index, index
        terms[t].active, terms[t].prior, terms[t].eta, terms[t].lambda_ = 1, prior, eta, 1.0
    ctx = Context(p)
    lib = ctx.lib
    d = ctx.image(img)
    native.check(lib.gvm_values_enqueue(ctx.ctx, d.ptr, flag_opt, int(normalize), 0,
                                        C.cast(terms, C.c_void_p), 3, 1), "values")
    fused = (C.c_double * 4)()
    native.check(lib.gvm_values_read(ctx.ctx, 4, fused), "read")
    d2 = ctx.image(img)
    native.check(lib.gvm_chi2_enqueue(ctx.ctx, d2.ptr, flag_opt, int(normalize), 4), "chi2")
    native.check(lib.gvm_entropy_value_enqueue(ctx.ctx, d2.ptr, 0, 0.001, -1.0, 5), "entropy")
    native.check(lib.gvm_tsv_value_enqueue(ctx.ctx, d2.ptr, 1, 6), "tsv")
    native.check(lib.gvm_laplacian_value_enqueue(ctx.ctx, d2.ptr, 0, 7), "laplacian")
    sep = (C.c_double * 8)()
    native.check(lib.gvm_values_read(ctx.ctx, 8, sep), "read")
    ctx.close()
    assert list(fused) == list(sep)[4:8], (list(fused), list(sep)[4:8])
    ref, _, _ = oracle.Oracle(p).chi2(img, flag_opt=flag_opt, normalize=normalize)
    assert abs(fused[0] - ref) <= CHI2_TOL * abs(ref), (fused[0], ref)
    assert all(v != 0.0 for v in fused[1:])
