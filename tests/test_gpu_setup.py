"""Set-up on the device (Hermitian flip, wavelengths, bilinear stencil and
off-grid rule, phase split, primary beam, beam/noise moments, noise image)
against the host set-up on the same raw rows."""
import ctypes as C

import numpy as np
import pytest

from gpuvmem_amd import native
from gpuvmem_amd.device import Context, DeviceBuffer
from gpuvmem_amd.problem import sim_config, small_problem

pytestmark = pytest.mark.gpu


def _raw_rows(p):
    """The generator's rows for problem p, in metres, before the flip."""
    fw = native.load_fw()
    n = fw.gvm_sim_count_visibilities(C.byref(p.sim))
    uvw = np.zeros((n, 3))
    vo = np.zeros((n, 2), np.float32)
    w = np.zeros(n, np.float32)
    assert fw.gvm_sim_generate(C.byref(p.sim), uvw.ctypes.data, vo.ctypes.data, w.ctypes.data,
                               None, None, None) == 0
    return uvw, vo, w


def test_device_setup_matches_host_setup():
    p = small_problem(npix=128, n_antennas=9, seed=5)
    uvw, vo, w = _raw_rows(p)
    w[::41] = 0.0
    # a few rows off the grid (|u| beyond N/2 cells)
    lam = float(np.float32(2.99792458e8) / np.float32(p.freq))
    uvw[::29, 0] = 3.2 * p.M / (p.M * abs(p.deltax) * np.pi / 180.0) * lam
    print(f"case: device set-up {p.M}^2 nvis={uvw.shape[0]}", flush=True)
    host = Context(p, add_data=False)
    u2, v2 = uvw.copy(), vo.copy()
    native.load_fw().gvm_setup_hermitian_lambda(u2.shape[0], u2.ctypes.data, v2.ctypes.data, p.freq)
    host.add(p.freq, u2, v2, w)
    dev = Context(p, add_data=False)
    bu, bv, bw = DeviceBuffer.from_array(uvw), DeviceBuffer.from_array(vo), DeviceBuffer.from_array(w)
    dev.add_dev(p.freq, bu, bv, bw, uvw.shape[0])
    rng = np.random.default_rng(2)
    img = p.initial_image()
    img[0] += (rng.random((p.N, p.M)) * 20).astype(np.float32)
    out = []
    for ctx in (host, dev):
        d = ctx.image(img)
        v = ctx.chi2(d)
        r = ctx.zeros(2 * p.M * p.N)
        ctx.dchi2(d, r)
        ctx.sync()
        out.append((v, ctx.residuals(), r.download(np.float32, (2, p.N, p.M))))
    (vh, rh, gh), (vd, rd, gd) = out
    assert abs(vh - vd) <= 1e-9 * abs(vh)
    assert np.abs(rh - rd).max() <= 1e-6 * np.abs(rh).max()
    assert np.linalg.norm(gh - gd) <= 1e-7 * np.linalg.norm(gh)
    # raw arrays are left unchanged
    assert np.array_equal(bu.download(np.float64, uvw.shape), uvw)
    assert np.array_equal(bv.download(np.float32, vo.shape), vo)
    host.close()
    dev.close()


def test_device_beam_noise_and_noise_image():
    lib, fw = native.load_hip(), native.load_fw()
    chans = []
    for f, seed in ((230e9, 1), (245e9, 2)):
        q = small_problem(npix=64, n_antennas=8, seed=seed, freq_hz=f)
        chans.append((np.float32(f), *_raw_rows(q)))
    nch = len(chans)
    freqs = (C.c_float * nch)(*[c[0] for c in chans])
    nv = (C.c_long * nch)(*[c[1].shape[0] for c in chans])
    hu = (C.c_void_p * nch)(*[c[1].ctypes.data for c in chans])
    hw = (C.c_void_p * nch)(*[c[3].ctypes.data for c in chans])
    bufs = [(DeviceBuffer.from_array(c[1]), DeviceBuffer.from_array(c[3])) for c in chans]
    du = (C.c_void_p * nch)(*[b[0].ptr for b in bufs])
    dw = (C.c_void_p * nch)(*[b[1].ptr for b in bufs])
    dx = -0.05 / 3600
    h = [C.c_double() for _ in range(6)]
    d = [C.c_double() for _ in range(6)]
    assert fw.gvm_setup_beam_noise(nch, freqs, nv, hu, hw, dx, *[C.byref(o) for o in h]) == 0
    assert lib.gvm_setup_beam_noise_dev(nch, freqs, nv, du, dw, dx, *[C.byref(o) for o in d]) == 0
    for a, b in zip(h, d):
        assert np.isclose(a.value, b.value, rtol=1e-12), (a.value, b.value)
    # the noise image, fg_scale and noise_cut
    args = (native.GVM_BEAM_GAUSSIAN, 12.0, 1.13, 1e-3, 230e9, 128, 128, dx, -dx, 64.0, 64.0,
            h[5].value, 10.0)
    nh, nd = np.zeros(128 * 128, np.float32), np.zeros(128 * 128, np.float32)
    fh, ch, fd, cd = C.c_float(), C.c_float(), C.c_float(), C.c_float()
    assert fw.gvm_setup_noise_image(*args, nh.ctypes.data, C.byref(fh), C.byref(ch)) == 0
    assert lib.gvm_setup_noise_image_dev(*args, nd.ctypes.data, C.byref(fd), C.byref(cd)) == 0
    finite = np.isfinite(nh)
    assert np.array_equal(finite, np.isfinite(nd))
    assert np.allclose(nd[finite], nh[finite], rtol=2e-7)
    assert np.isclose(fd.value, fh.value, rtol=2e-7) and np.isclose(cd.value, ch.value, rtol=2e-7)


def test_model_and_residual_output_in_observed_frame(tmp_path):
    """modelToHost: V_m = V_o - V_r conjugated back for the rows flipped at
    set-up; residuals likewise; and a FITS image written from the device."""
    import oracle
    from gpuvmem_amd.output import read_fits, write_fits

    p = small_problem(npix=64, n_antennas=7, seed=9)
    uvw, vo, w = _raw_rows(p)
    assert (uvw[:, 0] > 0).any() and (uvw[:, 0] <= 0).any()
    ctx = Context(p)  # the problem's rows are the flipped copies of these
    img = p.initial_image()
    img[0] += 5.0
    d = ctx.image(img)
    ctx.chi2(d)
    model = np.zeros((p.nvis, 2), np.float32)
    resid = np.zeros((p.nvis, 2), np.float32)
    native.check(ctx.lib.gvm_get_visibilities_out(ctx.ctx, 0, uvw.ctypes.data, model.ctypes.data,
                                                  resid.ctypes.data), "visibilities out")
    _, vr, _ = oracle.Oracle(p).chi2(img)
    flip = uvw[:, 0] > 0
    vm_ref = (p.vo[:, 0] + 1j * p.vo[:, 1]) - vr  # flipped frame
    vm_ref[flip] = np.conj(vm_ref[flip])
    vr_obs = vr.copy()
    vr_obs[flip] = np.conj(vr_obs[flip])
    scale = np.abs(vm_ref).max()
    assert np.abs(model[:, 0] + 1j * model[:, 1] - vm_ref).max() <= 1e-5 * scale
    assert np.abs(resid[:, 0] + 1j * resid[:, 1] - vr_obs).max() <= 1e-5 * scale
    # residuals + model reproduce the observed visibilities (the observed frame)
    assert np.allclose(model + resid, vo, atol=1e-5 * scale)
    path = str(tmp_path / "gpu.fits")
    write_fits(path, d, p, index=0, niter=3)
    _, data = read_fits(path)
    clipped = d.download(np.float32, (2, p.N, p.M))
    assert np.array_equal(data, clipped[0] * np.float32(p.fg_scale))
    ctx.close()
