"""Gridded-visibility mode on the GPU: the device gridder and the weighting
schemes against the oracle's restatement of the reference's single-thread
float arithmetic (bit for bit), and the objective of gridded slices with a
gridding correction image against the FP64 oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle
from gpuvmem_amd import native
from gpuvmem_amd.device import Context, DeviceBuffer
from gpuvmem_amd.gridding import (GriddingKernel, add_gridded_dev, apply_weighting_dev,
                                  configure_gridded, grid_visibilities_dev, set_gcf)
from gpuvmem_amd.problem import sim_config, small_problem

pytestmark = pytest.mark.gpu


def _raw_rows(cfg):
    fw = native.load_fw()
    n = fw.gvm_sim_count_visibilities(C.byref(cfg))
    uvw = np.zeros((n, 3))
    vo = np.zeros((n, 2), np.float32)
    w = np.zeros(n, np.float32)
    assert fw.gvm_sim_generate(C.byref(cfg), uvw.ctypes.data, vo.ctypes.data, w.ctypes.data,
                               None, None, None) == 0
    return uvw, vo, w


def _problem_rows(p, seed=0):
    uvw, vo, w = _raw_rows(p.sim)
    rng = np.random.default_rng(seed)
    w[rng.random(w.size) < 0.02] = 0.0
    lam = float(np.float32(2.99792458e8) / np.float32(p.freq))
    du = 1.0 / (p.M * abs(p.deltax) * np.pi / 180.0)
    uvw[::53, 0] = 0.7 * p.M * du * lam  # beyond the grid edge
    return uvw, vo, w


KERNELS = {
    "pillbox": lambda: GriddingKernel.pillbox(),
    "taps7x7": lambda: GriddingKernel.from_taps(
        np.random.default_rng(7).uniform(0.1, 1.0, (7, 7)).astype(np.float32)),
    "taps3x5": lambda: GriddingKernel.from_taps(
        np.random.default_rng(8).uniform(0.1, 1.0, (3, 5)).astype(np.float32)),
    "taps5x3": lambda: GriddingKernel.from_taps(
        np.random.default_rng(9).uniform(0.1, 1.0, (5, 3)).astype(np.float32)),
}


@pytest.mark.parametrize("kname", list(KERNELS))
def test_gridder_matches_oracle_bit_for_bit(kname):
    p = small_problem(npix=128, n_antennas=10, seed=11)
    uvw, vo, w = _problem_rows(p)
    k = KERNELS[kname]()
    print(f"case: gridder {kname} {p.M}^2 nvis={uvw.shape[0]}", flush=True)
    ctx = Context(p, add_data=False)
    bu, bv, bw = (DeviceBuffer.from_array(a) for a in (uvw, vo, w))
    gu, gv, gw = grid_visibilities_dev(ctx, k, p.freq, bu, bv, bw, uvw.shape[0])
    ou, ov, ow = oracle.grid_visibilities(p.M, p.N, p.deltax, p.deltay, p.freq, uvw, vo, w,
                                          k.taps, k.support_x, k.support_y)
    print(f"gridded rows: device {gu.shape[0]} oracle {ou.shape[0]}")
    assert gu.shape == ou.shape and gu.shape[0] > 100
    assert np.array_equal(gu, ou)
    assert np.array_equal(gv, ov)
    assert np.array_equal(gw, ow)
    # the raw rows are left unchanged
    assert np.array_equal(bw.download(np.float32), w)
    ctx.close()


def test_gridder_rejects_small_capacity_and_grids_empty_input():
    p = small_problem(npix=64, n_antennas=6, seed=12)
    uvw, vo, w = _problem_rows(p)
    print(f"case: gridder capacity/empty {p.M}^2 nvis={uvw.shape[0]}", flush=True)
    ctx = Context(p, add_data=False)
    bu, bv, bw = (DeviceBuffer.from_array(a) for a in (uvw, vo, w))
    k = GriddingKernel.pillbox()
    kc = k.as_c()
    out = [DeviceBuffer(24 * 8), DeviceBuffer(8 * 8), DeviceBuffer(4 * 8)]
    n = C.c_long()
    st = ctx.lib.gvm_grid_visibilities_dev(ctx.ctx, C.byref(kc), p.freq, uvw.shape[0], bu.ptr,
                                           bv.ptr, bw.ptr, out[0].ptr, out[1].ptr, out[2].ptr, 8,
                                           C.byref(n))
    assert st == native.GVM_ERR_ARG
    st = ctx.lib.gvm_grid_visibilities_dev(ctx.ctx, C.byref(kc), p.freq, 0, bu.ptr, bv.ptr,
                                           bw.ptr, out[0].ptr, out[1].ptr, out[2].ptr, 8,
                                           C.byref(n))
    assert st == native.GVM_OK and n.value == 0
    ctx.close()


def _cfg_at(p, f):
    cfg = sim_config(0, **{k: getattr(p.sim, k) for k in ("npix", "n_antennas", "ha_start_hours",
                                                          "ha_end_hours", "dump_seconds",
                                                          "cell_arcsec", "seed_layout")})
    cfg.freq_hz = float(f)
    return cfg


def _slices(p, freqs):
    out = []
    for f in freqs:
        uvw, vo, w = _raw_rows(_cfg_at(p, f))
        out.append((float(np.float32(f)), uvw, w))
    return out


TAPER = dict(amplitude=1.0, u_0=0.0, v_0=0.0, sigma_maj=4.0e5, sigma_min=2.5e5, bpa=0.6)


@pytest.mark.parametrize("scheme", ["natural", "uniform", "briggs", "radial"])
@pytest.mark.parametrize("taper", [None, TAPER])
def test_weighting_matches_oracle(scheme, taper):
    if scheme == "natural" and taper is None:
        pytest.skip("natural weighting without a taper changes nothing")
    p = small_problem(npix=128, n_antennas=10, seed=13)
    sl = _slices(p, [220e9, 230e9, 245e9])
    print(f"case: weighting {scheme} taper={taper is not None} {p.M}^2 "
          f"nvis={[s[1].shape[0] for s in sl]}", flush=True)
    ctx = Context(p, add_data=False)
    bufs = [(f, DeviceBuffer.from_array(u), DeviceBuffer.from_array(w), u.shape[0])
            for f, u, w in sl]
    apply_weighting_dev(ctx, scheme, bufs, robust=0.5, taper=taper)
    dev = [b[2].download(np.float32) for b in bufs]
    ref = oracle.weighting(scheme, sl, p.M, p.N, p.deltax, p.deltay, robust=0.5, taper=taper,
                           float_totals=False)
    for d, r in zip(dev, ref):
        if taper is None:
            assert np.array_equal(d, r)
        else:  # exp of the taper: device and host libm
            assert np.allclose(d, r, rtol=2e-7, atol=0)
    if scheme == "briggs":
        # against the reference's float totals: the difference is their rounding
        ref_f = oracle.weighting(scheme, sl, p.M, p.N, p.deltax, p.deltay, robust=0.5,
                                 taper=taper, float_totals=True)
        rel = max(float(np.max(np.abs(d - r) / np.maximum(np.abs(r), 1e-30)))
                  for d, r in zip(dev, ref_f))
        print(f"briggs vs float totals: max rel {rel:.2e}")
        assert rel < 1e-4
    ctx.close()


def test_weighting_rejects_bad_robust():
    p = small_problem(npix=64, n_antennas=6, seed=14)
    ctx = Context(p, add_data=False)
    with pytest.raises(native.GvmError):
        apply_weighting_dev(ctx, "briggs", [], robust=2.5)
    ctx.close()


@pytest.mark.parametrize("mode", ["f16x3", "bf16x3", "direct"])
def test_gridded_objective_with_gcf_matches_oracle(mode):
    p = small_problem(npix=128, n_antennas=10, seed=15)
    uvw, vo, w = _problem_rows(p, seed=1)
    k = KERNELS["taps7x7"]()
    print(f"case: gridded objective {mode} {p.M}^2 nvis={uvw.shape[0]}", flush=True)
    rng = np.random.default_rng(5)
    gcf = rng.uniform(0.5, 1.5, (p.N, p.M)).astype(np.float32)
    ctx = Context(p, adjoint=mode, add_data=False)
    bu, bv, bw = (DeviceBuffer.from_array(a) for a in (uvw, vo, w))
    apply_weighting_dev(ctx, "uniform", [(p.freq, bu, bw, uvw.shape[0])])
    s, ng = add_gridded_dev(ctx, k, p.freq, bu, bv, bw, uvw.shape[0])
    set_gcf(ctx, gcf)
    # the oracle's pipeline: uniform weights, gridding, Hermitian flip
    (w2,) = oracle.weighting("uniform", [(p.freq, uvw, w)], p.M, p.N, p.deltax, p.deltay)
    gu, gv, gw = oracle.grid_visibilities(p.M, p.N, p.deltax, p.deltay, p.freq, uvw, vo, w2,
                                          k.taps, k.support_x, k.support_y)
    assert gu.shape[0] == ng
    oracle.lib().orc_hermitian_to_lambda(gu.shape[0], gu.ctypes.data, gv.ctypes.data, p.freq)
    orc = oracle.Oracle(p, slices=[(p.freq, gu, gv, gw)], gcf=gcf)
    img = p.initial_image()
    img[0] += (rng.random((p.N, p.M)) * 20).astype(np.float32)
    for normalize in (False, True):
        d = ctx.image(img)
        v = ctx.chi2(d, normalize=normalize)
        ref, vr, clipped = orc.chi2(img, normalize=normalize)
        assert abs(v - ref) <= 1e-6 * abs(ref), (v, ref)
        r = ctx.zeros(2 * p.M * p.N)
        ctx.dchi2(d, r, normalize=normalize)
        ctx.sync()
        g = r.download(np.float32, (2, p.N, p.M))
        gref = orc.dchi2(clipped, vr, normalize=normalize)
        err = np.linalg.norm(g[0] - gref[0]) / np.linalg.norm(gref[0])
        print(f"normalize={normalize}: chi2 rel {abs(v - ref) / abs(ref):.2e} grad rel L2 {err:.2e}")
        assert err < (3e-5 if mode == "bf16x3" else 1e-5)
    ctx.close()


def test_unit_gcf_changes_nothing():
    p = small_problem(npix=64, n_antennas=8, seed=16)
    print(f"case: unit gcf {p.M}^2 nvis={p.nvis}", flush=True)
    img = p.initial_image()
    img[0] += np.random.default_rng(1).random((p.N, p.M)).astype(np.float32)
    out = []
    for gcf in (None, np.ones((p.N, p.M), np.float32)):
        ctx = Context(p)
        set_gcf(ctx, gcf)
        d = ctx.image(img)
        v = ctx.chi2(d)
        r = ctx.zeros(2 * p.M * p.N)
        ctx.dchi2(d, r)
        ctx.sync()
        out.append((v, r.download(np.float32)))
        ctx.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("kind,params", [("pillbox", {}), ("gaussian", {"w": 3.0}),
                                         ("pswf", {})])
def test_configure_gridded_with_built_kernels(kind, params):
    """Briggs weighting, gridding with a kernel of the reference's family and
    its correction image, over two channels, against the oracle."""
    p = small_problem(npix=128, n_antennas=10, seed=17)
    k = GriddingKernel.build(kind, p.M, p.N, p.deltax, p.deltay, **params)
    sl = _slices(p, [226e9, 234e9])
    vos = [_raw_rows(_cfg_at(p, f))[1] for f, _, _ in sl]
    print(f"case: configure gridded {kind} {p.M}^2 nvis={[s[1].shape[0] for s in sl]}", flush=True)
    ctx = Context(p, adjoint="f16x3", add_data=False)
    bufs = [(f, DeviceBuffer.from_array(u), DeviceBuffer.from_array(v), DeviceBuffer.from_array(w),
             u.shape[0]) for (f, u, w), v in zip(sl, vos)]
    counts = configure_gridded(ctx, k, bufs, scheme="briggs", robust=0.0)
    ws = oracle.weighting("briggs", sl, p.M, p.N, p.deltax, p.deltay, robust=0.0,
                          float_totals=False)
    oslices = []
    for (f, u, _), v, w in zip(sl, vos, ws):
        gu, gv, gw = oracle.grid_visibilities(p.M, p.N, p.deltax, p.deltay, f, u, v, w, k.taps,
                                              k.support_x, k.support_y)
        oracle.lib().orc_hermitian_to_lambda(gu.shape[0], gu.ctypes.data, gv.ctypes.data, f)
        oslices.append((f, gu, gv, gw))
    assert counts == [s[1].shape[0] for s in oslices]
    orc = oracle.Oracle(p, slices=oslices, gcf=k.gcf)
    img = p.initial_image()
    img[0] += (np.random.default_rng(3).random((p.N, p.M)) * 10).astype(np.float32)
    d = ctx.image(img)
    v = ctx.chi2(d)
    ref, vr, clipped = orc.chi2(img)
    r = ctx.zeros(2 * p.M * p.N)
    ctx.dchi2(d, r)
    ctx.sync()
    g = r.download(np.float32, (2, p.N, p.M))
    gref = orc.dchi2(clipped, vr)
    err = np.linalg.norm(g[0] - gref[0]) / np.linalg.norm(gref[0])
    print(f"chi2 rel {abs(v - ref) / abs(ref):.2e} grad rel L2 {err:.2e} rows {counts}")
    assert abs(v - ref) <= 1e-6 * abs(ref)
    assert err < 1e-5
    ctx.close()
