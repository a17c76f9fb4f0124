"""Host-side checks that need no GPU: the C ABI libraries load and export every
declared entry point, host-forcible errors are reported as status codes, the
set-up and generator agree with the oracle / closed forms, and the factories
hold the reference's ids."""
import ctypes as C
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from gpuvmem_amd import native
from gpuvmem_amd.problem import make_problem, sim_config, small_problem

ROOT = Path(__file__).resolve().parents[1]


def _declared(header: Path) -> set:
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(gvm_[a-z0-9_]+)\s*\(", text))


def _exported(lib: Path) -> set:
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_hip_library_exports_every_declared_symbol():
    decl = _declared(ROOT / "include" / "gvm_hip.h") | _declared(ROOT / "include" / "gvm_nn.h")
    assert len(decl) >= 35
    exp = _exported(native.LIB_DIR / "libgvm_hip.so")
    assert decl <= exp, sorted(decl - exp)
    # and the bindings know all of them
    assert decl <= set(native.HIP_API), sorted(decl - set(native.HIP_API))


def test_framework_library_exports_every_declared_symbol():
    exp = _exported(native.LIB_DIR / "libgvm.so")
    for h in ("fw.h", "setup.h", "simdata.h", "fitsio.h", "ckernel.h"):
        decl = _declared(ROOT / "include" / "gvm" / h)
        assert decl, h
        assert decl <= exp, (h, sorted(decl - exp))


def test_libraries_load_without_gpu():
    lib = native.load_hip()
    native.load_fw()
    n = C.c_int(-1)
    st = lib.gvm_device_info(None, 0, C.byref(n), None)
    assert st in (native.GVM_OK, 5)


def test_create_reports_errors_as_status_codes():
    lib = native.load_hip()
    p = small_problem()
    noise = np.ascontiguousarray(p.noise, np.float32)
    ctx = C.c_void_p()
    g = p.geometry()
    g.M = 100  # not square / not a power of two
    assert lib.gvm_create(C.byref(g), noise.ctypes.data, C.byref(ctx)) == 1
    assert b"M must equal N" in lib.gvm_last_error()
    g = p.geometry()
    g.x0_ref = float(p.M + 3)  # pointing centre outside the image (src/mfs.cu:708-731)
    assert lib.gvm_create(C.byref(g), noise.ctypes.data, C.byref(ctx)) == 1
    g = p.geometry()
    g.image_count = 1
    assert lib.gvm_create(C.byref(g), noise.ctypes.data, C.byref(ctx)) == 1
    assert lib.gvm_create(None, None, C.byref(ctx)) == 1
    # a valid geometry on a machine without a gfx950 device
    g = p.geometry()
    st = lib.gvm_create(C.byref(g), noise.ctypes.data, C.byref(ctx))
    if st == 0:
        lib.gvm_destroy(ctx)
    else:
        assert st == 5
    # calls on a null context are rejected, not crashed
    out = C.c_double()
    assert lib.gvm_chi2(None, None, 0, 0, C.byref(out)) == 1
    assert lib.gvm_vec_dot(None, None, None, 4, C.byref(out)) == 1


def test_primary_beam_table_matches_oracle():
    lib = native.load_hip()
    p = small_problem(npix=64)
    pb = np.zeros((p.N, p.M), np.float32)
    st = lib.gvm_primary_beam(p.beam_kind, p.antenna_diameter, p.pb_factor, p.pb_cutoff, p.freq,
                              p.M, p.N, p.x0, p.y0, p.deltax, p.deltay, pb.ctypes.data)
    assert st == 0
    ref = oracle.Oracle(p).pb()
    assert np.allclose(pb, ref, rtol=1e-7, atol=1e-30)
    assert pb[int(p.y0), int(p.x0)] == 1.0
    # Airy disk: 1 at the centre, first null at pb_factor lambda / D scaled by RZ
    pa = np.zeros((p.N, p.M), np.float32)
    assert lib.gvm_primary_beam(0, 12.0, 1.13, 1.0, p.freq, p.M, p.N, p.x0, p.y0, p.deltax,
                                p.deltay, pa.ctypes.data) == 0
    b = oracle.OBeam(0, 12.0, 1.13, 1.0)
    ref_a = np.array([[oracle.lib().orc_attenuation(C.byref(b), p.freq, i, j, p.x0, p.y0, p.deltax,
                                                    p.deltay) for j in range(p.M)] for i in range(p.N)])
    assert np.allclose(pa, ref_a, rtol=1e-6)
    assert lib.gvm_primary_beam(7, 12.0, 1.13, 1.0, p.freq, p.M, p.N, p.x0, p.y0, p.deltax,
                                p.deltay, pa.ctypes.data) == 1


def test_setup_matches_oracle_restatement():
    fw = native.load_fw()
    cfg = sim_config(0, n_antennas=6, dump_seconds=120.0)
    n = fw.gvm_sim_count_visibilities(C.byref(cfg))
    uvw = np.zeros((n, 3))
    vo = np.zeros((n, 2), np.float32)
    w = np.zeros(n, np.float32)
    assert fw.gvm_sim_generate(C.byref(cfg), uvw.ctypes.data, vo.ctypes.data, w.ctypes.data, None,
                               None, None) == 0
    freq = np.float32(cfg.freq_hz)
    out = [C.c_double() for _ in range(6)]
    fq = (C.c_float * 1)(freq)
    nv = (C.c_long * 1)(n)
    up = (C.c_void_p * 1)(uvw.ctypes.data)
    wp = (C.c_void_p * 1)(w.ctypes.data)
    assert fw.gvm_setup_beam_noise(1, fq, nv, up, wp, -0.05 / 3600, *[C.byref(o) for o in out]) == 0
    ch = (oracle.OChan * 1)(oracle.OChan(freq, n, uvw.ctypes.data, vo.ctypes.data, w.ctypes.data))
    ref = [C.c_double() for _ in range(5)]
    oracle.lib().orc_beam_and_noise(ch, 1, -0.05 / 3600, *[C.addressof(r) for r in ref])
    for a, b in zip(out[:5], ref):
        assert np.isclose(a.value, b.value, rtol=1e-12)
    njp = oracle.lib().orc_noise_jypix(ref[4].value, ref[0].value, ref[1].value, -0.05 / 3600)
    assert np.isclose(out[5].value, njp, rtol=1e-12)
    # zero weights are rejected ("The sum of the visibility weights cannot be zero")
    w0 = np.zeros(n, np.float32)
    wp0 = (C.c_void_p * 1)(w0.ctypes.data)
    assert fw.gvm_setup_beam_noise(1, fq, nv, up, wp0, -0.05 / 3600, *[C.byref(o) for o in out]) == 1
    # Hermitian flip and wavelengths
    a, va = uvw.copy(), vo.copy()
    b, vb = uvw.copy(), vo.copy()
    fw.gvm_setup_hermitian_lambda(n, a.ctypes.data, va.ctypes.data, freq)
    oracle.lib().orc_hermitian_to_lambda(n, b.ctypes.data, vb.ctypes.data, freq)
    assert np.array_equal(a, b) and np.array_equal(va, vb)
    assert np.all(a[:, 0] <= 0)


def test_generator_counts_and_determinism():
    fw = native.load_fw()
    for which, n, npix in ((0, 36000, 256), (1, 1764000, 1024), (2, 4354560, 2048),
                           (3, 299880, 512), (4, 1029000, 512)):
        cfg = sim_config(which)
        assert fw.gvm_sim_count_visibilities(C.byref(cfg)) == n
        assert cfg.npix == npix
    p1 = small_problem(seed=5)
    p2 = small_problem(seed=5)
    assert np.array_equal(p1.uvw, p2.uvw) and np.array_equal(p1.vo, p2.vo)
    p3 = small_problem(seed=5, seed_noise=99)
    assert np.array_equal(p1.uvw, p3.uvw) and not np.array_equal(p1.vo, p3.vo)
    assert np.all(p1.weight == np.float32(1e6))
    # noise statistics: sigma 1 mJy per component
    d = (p1.vo - p3.vo).ravel()
    assert 0.8e-3 < d.std() / np.sqrt(2) < 1.2e-3


def test_generator_spectral_index_and_flagging():
    """Multi-frequency generator: fluxes scale as (nu / nu0)^alpha (a point
    source at the phase centre, no noise: every visibility equals its flux),
    flagging packs the surviving rows in order, and the single-channel
    workloads are unchanged by the new fields."""
    fw = native.load_fw()
    cfg = sim_config(3, npix=64, n_antennas=5, ha_start_hours=-0.05, ha_end_hours=0.05,
                     n_gaussians=0, n_points=1, position_radius_frac=0.0, noise_sigma=1e-9,
                     point_flux_min=0.5, point_flux_max=0.5, alpha_min=2.0, alpha_max=2.0)
    n = fw.gvm_sim_count_visibilities(C.byref(cfg))
    rows = {}
    for f, frac in ((225e9, 0.0), (240e9, 0.0), (240e9, 0.3)):
        cfg.freq_hz, cfg.flag_fraction = f, frac
        uvw = np.zeros((n, 3))
        vo = np.zeros((n, 2), np.float32)
        w = np.zeros(n, np.float32)
        nk = C.c_long()
        assert fw.gvm_sim_generate_flagged(C.byref(cfg), uvw.ctypes.data, vo.ctypes.data,
                                           w.ctypes.data, None, None, None, C.byref(nk)) == 0
        rows[(f, frac)] = (uvw[:nk.value], vo[:nk.value])
    # centre point: flux * PB(centre) = flux, times (nu/nu0)^alpha
    assert np.allclose(rows[(225e9, 0.0)][1][:, 0], 0.5, rtol=1e-5)
    assert np.allclose(rows[(240e9, 0.0)][1][:, 0], 0.5 * (240 / 225) ** 2, rtol=1e-5)
    kept = rows[(240e9, 0.3)][0]
    assert 0.5 * n < kept.shape[0] < 0.9 * n
    full = rows[(240e9, 0.0)][0]
    idx = [np.flatnonzero((full == r).all(axis=1))[0] for r in kept[:20]]
    assert idx == sorted(idx)  # surviving rows keep their order
    cfg.flag_fraction = 1.0
    assert fw.gvm_sim_generate_flagged(C.byref(cfg), uvw.ctypes.data, vo.ctypes.data,
                                       w.ctypes.data, None, None, None, C.byref(nk)) == 1


def test_problem_setup_invariants():
    p = make_problem(0)
    assert p.nvis == 36000 and p.M == p.N == 256
    assert np.all(p.uvw[:, 0] <= 0)
    assert np.isclose(p.fg_scale, p.beam_info["noise_jypix"], rtol=1e-6)
    assert np.isclose(p.noise_cut, 10 * p.fg_scale, rtol=1e-6)
    assert p.noise[int(p.y0), int(p.x0)] == np.float32(p.fg_scale)


def test_factories_hold_reference_ids():
    fw = native.load_fw()
    for name in ("Chi2", "Entropy", "TotalSquaredVariation", "Laplacian"):
        assert fw.gvm_fw_registered(b"Fi", name.encode()) == 1, name
    for name in ("CG-FRPRMN", "CG-LBFGS"):
        assert fw.gvm_fw_registered(b"Optimizer", name.encode()) == 1, name
    assert fw.gvm_fw_registered(b"ObjectiveFunction", b"ObjectiveFunction") == 1
    assert fw.gvm_fw_registered(b"Fi", b"L1-Norm") == 0  # not on this path
    assert fw.gvm_fw_registered(b"Optimizer", b"nope") == 0


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_oracle_is_not_reachable_from_the_product():
    """The product package never imports, links or opens the oracle."""
    for f in list((ROOT / "gpuvmem_amd").rglob("*.py")) + list((ROOT / "gpuvmem_amd").rglob("*.cpp")) \
            + list((ROOT / "gpuvmem_amd").rglob("*.hip")):
        t = f.read_text()
        assert not re.search(r"^\s*(import|from)\s+oracle|liboracle|gvm_oracle", t, re.M), f
    ldd = subprocess.run(["ldd", str(native.LIB_DIR / "libgvm.so")], capture_output=True, text=True).stdout
    assert "oracle" not in ldd


def test_fits_writer_round_trip(tmp_path):
    """The FITS writer: fixed-format cards, 2880-byte records, big-endian
    BITPIX -32 data equal to image * fg_scale, the keys OCopyFITS sets."""
    from gpuvmem_amd.output import read_fits, write_fits

    p = small_problem(npix=64)
    rng = np.random.default_rng(4)
    img = rng.standard_normal((2, p.N, p.M)).astype(np.float32)
    path = str(tmp_path / "mem.fits")
    write_fits(path, img, p, index=0, niter=37)
    raw = open(path, "rb").read()
    assert len(raw) % 2880 == 0
    assert raw[:30].decode() == "SIMPLE  =                    T"
    assert raw[80:110].decode() == "BITPIX  =                  -32"
    hdr, data = read_fits(path)
    assert (hdr["NAXIS1"], hdr["NAXIS2"], hdr["NITER"], hdr["BUNIT"]) == (p.M, p.N, 37, "JY/PIXEL")
    assert hdr["CDELT1"] == pytest.approx(p.deltax, rel=1e-12) and hdr["CRPIX1"] == p.x0 + 1
    assert hdr["EQUINOX"] == 2000.0 and hdr["RADESYS"] == "ICRS"
    assert np.array_equal(data, img[0] * np.float32(p.fg_scale))
    write_fits(path, img, p, index=1)  # alpha: unscaled, no unit
    hdr, data = read_fits(path)
    assert np.array_equal(data, img[1]) and hdr["BUNIT"] == ""
    with pytest.raises(OSError):
        write_fits(str(tmp_path / "no" / "dir" / "x.fits"), img, p)


def _kernel(kind, M=64, dx=-1e-5, with_gcf=True, **kw):
    from gpuvmem_amd.gridding import GriddingKernel
    return GriddingKernel.build(kind, M, M, dx, -dx, with_gcf=with_gcf, **kw)


def test_ckernel_pillbox_and_gaussian_closed_forms():
    k = _kernel("pillbox")
    assert k.taps.shape == (1, 1) and k.taps[0, 0] == 1.0 and k.support_x == 0
    assert np.all(k.gcf == 1.0)
    k = _kernel("pillbox", m=5, n=5)
    assert np.all(k.taps == 1.0) and k.support_x == k.support_y == 2
    # Gaussian2D: exp(-(|x|/(w s))^a - (|y|/(w s))^a) inside |x|, |y| < w s
    for w, a in ((3.0, 2.0), (2.5, 1.0)):
        k = _kernel("gaussian", w=w, alpha=a)
        d = np.abs(np.arange(7) - 3)
        f = np.where(d < w, np.exp(-(d / w) ** a), 0.0)
        ref = np.outer(f, f)
        assert np.allclose(k.taps, ref, rtol=2e-6, atol=0)
    # the default (w = 1) keeps the centre tap only
    k = _kernel("gaussian")
    assert k.taps[3, 3] == 1.0 and np.count_nonzero(k.taps) == 1
    # its correction image: the Gaussian at pi x with width 2 M cells
    M = 64
    d = np.abs(np.arange(M) - M // 2)
    g = np.exp(-(np.pi * d / (2 * M)) ** 2)
    assert np.allclose(k.gcf, np.outer(g, g), rtol=2e-6)


def test_ckernel_sinc_and_gaussian_sinc_closed_forms():
    # Sinc2D::buildKernel hands (x, y) to sinc2D's (x, x0) slots: the tap
    # is sinc(|x - y| / (w s)), a band along the diagonal
    k = _kernel("sinc", with_gcf=False, w=3.0)
    i, j = np.indices((7, 7))
    ref = np.where(np.abs(j - i) < 3, np.sinc((j - i) / 3.0), 0.0)
    assert np.allclose(k.taps, ref, rtol=2e-6, atol=1e-7)
    # GaussianSinc2D (w = 2.52, w2 = 1.55, alpha = 2): nonzero for |x|, |y| <= 1 cell
    k = _kernel("gaussian_sinc", with_gcf=False)
    d = np.abs(np.arange(7) - 3).astype(float)
    f = np.where(d < 1.55, np.exp(-(d / 2.52) ** 2) * np.sinc(d / 1.55), 0.0)
    assert np.allclose(k.taps, np.outer(f, f), rtol=2e-6, atol=1e-7)
    # neither defines a correction image in the reference
    from gpuvmem_amd.gridding import GriddingKernel
    for kind in ("sinc", "gaussian_sinc"):
        with pytest.raises(ValueError):
            GriddingKernel.build(kind, 64, 64, -1e-5, 1e-5)


def test_ckernel_pswf_properties():
    # a long 1-D profile: m = 201, w = 100 gives nu = |j - 100| / 100
    k = _kernel("pswf", with_gcf=False, m=201, n=201, w=100.0)
    row = k.taps[100].astype(np.float64)
    nu = np.abs(np.arange(201) - 100) / 100.0
    assert row[100] == 1.0 and np.allclose(row, row[::-1])
    psi = row[101:200] / (1 - nu[101:200] ** 2)
    # Schwab's function is normalised to 1 at nu = 0 and smooth across the
    # break between its two rational pieces at nu = 0.75
    assert abs(psi[0] - 1.0) < 1e-3
    at = 74  # nu = 0.75
    second = psi[at - 1] - 2 * psi[at] + psi[at + 1]
    assert abs(second) < 5e-3 * psi[at]
    assert np.all(np.diff(psi) < 0)
    assert row[200] == 0.0 and 0 < row[199] < 1e-2
    # the correction image is the reciprocal of the separable kernel with w = M
    k = _kernel("pswf")
    g = k.gcf.astype(np.float64)
    assert g[32, 32] == 1.0
    assert np.allclose(g, np.outer(g[:, 32], g[32, :]), rtol=1e-5)
    assert np.all(np.diff(g[32, 32:]) > 0)
