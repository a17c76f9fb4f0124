"""Imaging problems: synthetic observation + the reference's set-up step.

A :class:`Problem` holds what the hot path consumes after the reference's
``MFS::configure`` / ``MFS::setDevice`` (src/mfs.cu:79-943): uvw in
wavelengths after the Hermitian flip, Vo, weights, the noise image (mask),
fg_scale, noise_cut, the beam, and the image geometry.  Data come from the
native generator (``gvm_sim_generate``) and the native set-up helpers
(``gvm_setup_*``); nothing here touches the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import native

LIGHTSPEED = 2.99792458e8


@dataclass
class Problem:
    M: int
    N: int
    deltax: float  # degrees (CDELT1, negative)
    deltay: float  # degrees (CDELT2)
    x0: float
    y0: float
    freq: float
    nu_0: float
    uvw: np.ndarray  # (nvis, 3) float64, wavelengths, u <= 0
    vo: np.ndarray  # (nvis, 2) float32
    weight: np.ndarray  # (nvis,) float32
    noise: np.ndarray  # (N, M) float32
    fg_scale: float
    noise_cut: float
    beam_kind: int
    antenna_diameter: float
    pb_factor: float
    pb_cutoff: float
    minpix: float = 0.001
    eta: float = -1.0
    threshold: float = 0.0
    sky: Optional[np.ndarray] = None
    sim: Optional[native.SimConfig] = None
    beam_info: dict = field(default_factory=dict)
    # multi-frequency: every (freq, uvw, vo, weight) channel x stokes slice;
    # None means the single slice held in freq / uvw / vo / weight
    slices: Optional[list] = None

    @property
    def nvis(self) -> int:
        """Visibilities of the first slice (the only one for one channel)."""
        return int(self.uvw.shape[0])

    @property
    def nvis_total(self) -> int:
        return sum(int(u.shape[0]) for _, u, _, _ in self.all_slices())

    def all_slices(self) -> list:
        if self.slices is not None:
            return self.slices
        return [(self.freq, self.uvw, self.vo, self.weight)]

    def geometry(self) -> native.Geometry:
        g = native.Geometry()
        g.M, g.N, g.image_count = self.M, self.N, 2
        g.deltax, g.deltay = self.deltax, self.deltay
        g.x0_ref = g.x0_phs = self.x0
        g.y0_ref = g.y0_phs = self.y0
        g.nu_0 = self.nu_0
        g.fg_scale = self.fg_scale
        g.noise_cut = self.noise_cut
        g.minpix = self.minpix
        g.eta = self.eta
        g.threshold = self.threshold
        g.beam_kind = self.beam_kind
        g.antenna_diameter = self.antenna_diameter
        g.pb_factor = self.pb_factor
        g.pb_cutoff = self.pb_cutoff
        return g

    def initial_image(self, value: Optional[float] = None) -> np.ndarray:
        """The reference's starting image (src/mfs.cu:737-745): I_nu_0 =
        initial_values[0] everywhere, alpha = 0.  Shape (2, N, M) float32."""
        img = np.zeros((2, self.N, self.M), np.float32)
        img[0] = self.minpix if value is None else value
        return img


def sim_config(which: int = 0, **overrides) -> native.SimConfig:
    fw = native.load_fw()
    cfg = native.SimConfig()
    fw.gvm_sim_default_config(which, C.byref(cfg))
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise KeyError(f"unknown simulation field {k}")
        setattr(cfg, k, v)
    return cfg


def make_problem(which: int = 0, noise_cut_factor: float = 10.0, **overrides) -> Problem:
    """Generates workload ``which`` (0: 256^2 / 36,000 vis; 1: 1024^2 /
    1,764,000 vis; 2: 2048^2 / 4,354,560 vis; 4: 512^2 / 1,029,000 vis, the
    warm start's) with optional field overrides and runs the set-up."""
    fw = native.load_fw()
    cfg = sim_config(which, **overrides)
    nvis = fw.gvm_sim_count_visibilities(C.byref(cfg))
    if nvis <= 0:
        raise ValueError("empty observation")
    uvw_m = np.zeros((nvis, 3), np.float64)
    vo = np.zeros((nvis, 2), np.float32)
    weight = np.zeros(nvis, np.float32)
    sky = np.zeros((cfg.npix, cfg.npix), np.float32)
    comps = (native.SimComponent * (cfg.n_gaussians + cfg.n_points))()
    ncomp = C.c_int(0)
    rc = fw.gvm_sim_generate(C.byref(cfg), uvw_m.ctypes.data, vo.ctypes.data, weight.ctypes.data,
                             sky.ctypes.data, comps, C.byref(ncomp))
    if rc != 0:
        raise RuntimeError("gvm_sim_generate failed")

    npix = cfg.npix
    deltax = -cfg.cell_arcsec / 3600.0
    deltay = cfg.cell_arcsec / 3600.0
    x0 = y0 = float(npix // 2)  # CRPIX = npix/2 + 1, phase centre = pointing = image centre
    freq = float(np.float32(cfg.freq_hz))

    # calculateNoiseAndBeam + noise_jypix (metres, before the Hermitian flip)
    freqs = (C.c_float * 1)(freq)
    nv = (C.c_long * 1)(nvis)
    uptr = (C.c_void_p * 1)(uvw_m.ctypes.data)
    wptr = (C.c_void_p * 1)(weight.ctypes.data)
    out = [C.c_double() for _ in range(6)]
    rc = fw.gvm_setup_beam_noise(1, freqs, nv, uptr, wptr, deltax, *[C.byref(o) for o in out])
    native.check(rc, "gvm_setup_beam_noise")
    bmaj, bmin, bpa, sumw, vis_noise, noise_jypix = (o.value for o in out)

    lam_max = float(np.float32(LIGHTSPEED) / np.float32(freq))
    pb_cutoff = fw.gvm_setup_pb_cutoff(cfg.pb_factor, lam_max, cfg.dish_diameter_m)
    noise = np.zeros((npix, npix), np.float32)
    fg = C.c_float()
    ncut = C.c_float()
    rc = fw.gvm_setup_noise_image(native.GVM_BEAM_GAUSSIAN, cfg.dish_diameter_m, cfg.pb_factor,
                                  pb_cutoff, freq, npix, npix, deltax, deltay, x0, y0, noise_jypix,
                                  noise_cut_factor, noise.ctypes.data, C.byref(fg), C.byref(ncut))
    native.check(rc, "gvm_setup_noise_image")

    uvw = uvw_m.copy()
    fw.gvm_setup_hermitian_lambda(nvis, uvw.ctypes.data, vo.ctypes.data, freq)

    return Problem(
        M=npix, N=npix, deltax=deltax, deltay=deltay, x0=x0, y0=y0, freq=freq, nu_0=freq,
        uvw=uvw, vo=vo, weight=weight, noise=noise, fg_scale=float(fg.value),
        noise_cut=float(ncut.value), beam_kind=native.GVM_BEAM_GAUSSIAN,
        antenna_diameter=float(np.float32(cfg.dish_diameter_m)),
        pb_factor=float(np.float32(cfg.pb_factor)), pb_cutoff=float(np.float32(pb_cutoff)),
        sky=sky, sim=cfg,
        beam_info=dict(bmaj_deg=bmaj, bmin_deg=bmin, bpa_deg=bpa, sum_weights=sumw,
                       vis_noise=vis_noise, noise_jypix=noise_jypix),
    )


def small_problem(npix: int = 64, n_antennas: int = 8, seed: int = 0, **overrides) -> Problem:
    """A problem small enough for FP64 direct sums in tests."""
    kw = dict(npix=npix, n_antennas=n_antennas, ha_start_hours=-0.5, ha_end_hours=0.5,
              dump_seconds=60.0, cell_arcsec=0.05, n_gaussians=3, n_points=2,
              position_radius_frac=0.25, seed_layout=seed, seed_sky=seed + 1,
              seed_noise=seed + 2, sigma_min_pix=1.5, sigma_max_pix=4.0)
    kw.update(overrides)
    return make_problem(0, **kw)


def mfs_frequencies(n_spw: int = 4, n_chan: int = 64, f_lo: float = 210e9,
                    f_hi: float = 240e9) -> np.ndarray:
    """Channel centres of n_spw contiguous spectral windows of n_chan
    channels spanning [f_lo, f_hi], as float32 values (the reference keeps
    frequencies in float)."""
    n = n_spw * n_chan
    width = (f_hi - f_lo) / n
    return np.array([np.float32(f_lo + (k + 0.5) * width) for k in range(n)], np.float64)


def make_mfs_problem(n_spw: int = 4, n_chan: int = 64, flag_max: float = 0.4,
                     noise_cut_factor: float = 10.0, seed: int = 300, **overrides) -> Problem:
    """The multi-frequency workload: 512^2 I_nu_0 + alpha, n_spw x n_chan
    channels over 210-240 GHz, 299,880 rows per channel before flagging, a
    per-channel flag fraction U(0, flag_max) (uneven counts), one sky with
    spectral indices U(-1, 3) about nu_0 = 225 GHz.  The set-up follows
    MFS::configure / setDevice over all channels: beam and noise from every
    channel's weights (calculateNoiseAndBeam), the noise image from the
    primary beam at nu_0 (src/mfs.cu:849-916), Hermitian flip and
    wavelengths per channel."""
    fw = native.load_fw()
    base = sim_config(3, **overrides)
    freqs = mfs_frequencies(n_spw, n_chan)
    rng = np.random.default_rng(seed)
    flags = rng.uniform(0.0, flag_max, freqs.size)
    nmax = fw.gvm_sim_count_visibilities(C.byref(base))
    if nmax <= 0:
        raise ValueError("empty observation")
    npix = base.npix
    sky = np.zeros((npix, npix), np.float32)
    chans = []
    for k, f in enumerate(freqs):
        cfg = sim_config(3, **overrides)
        cfg.freq_hz = float(f)
        cfg.flag_fraction = float(flags[k])
        cfg.seed_noise = base.seed_noise + 1000 * (k + 1)
        cfg.seed_flag = base.seed_flag + 1000 * (k + 1)
        uvw = np.zeros((nmax, 3), np.float64)
        vo = np.zeros((nmax, 2), np.float32)
        w = np.zeros(nmax, np.float32)
        nk = C.c_long()
        rc = fw.gvm_sim_generate_flagged(C.byref(cfg), uvw.ctypes.data, vo.ctypes.data,
                                         w.ctypes.data, sky.ctypes.data if k == 0 else None,
                                         None, None, C.byref(nk))
        if rc != 0:
            raise RuntimeError("gvm_sim_generate_flagged failed")
        n = nk.value
        chans.append((float(np.float32(f)), uvw[:n].copy(), vo[:n].copy(), w[:n].copy()))

    deltax = -base.cell_arcsec / 3600.0
    deltay = base.cell_arcsec / 3600.0
    x0 = y0 = float(npix // 2)
    nu0 = float(np.float32(base.nu0_hz))
    nch = len(chans)
    freqs_c = (C.c_float * nch)(*[c[0] for c in chans])
    nv = (C.c_long * nch)(*[c[1].shape[0] for c in chans])
    uptr = (C.c_void_p * nch)(*[c[1].ctypes.data for c in chans])
    wptr = (C.c_void_p * nch)(*[c[3].ctypes.data for c in chans])
    out = [C.c_double() for _ in range(6)]
    native.check(fw.gvm_setup_beam_noise(nch, freqs_c, nv, uptr, wptr, deltax,
                                         *[C.byref(o) for o in out]), "gvm_setup_beam_noise")
    bmaj, bmin, bpa, sumw, vis_noise, noise_jypix = (o.value for o in out)
    lam_max = float(np.float32(LIGHTSPEED) / np.float32(min(c[0] for c in chans)))
    pb_cutoff = fw.gvm_setup_pb_cutoff(base.pb_factor, lam_max, base.dish_diameter_m)
    noise = np.zeros((npix, npix), np.float32)
    fg, ncut = C.c_float(), C.c_float()
    native.check(fw.gvm_setup_noise_image(native.GVM_BEAM_GAUSSIAN, base.dish_diameter_m,
                                          base.pb_factor, pb_cutoff, nu0, npix, npix, deltax, deltay,
                                          x0, y0, noise_jypix, noise_cut_factor, noise.ctypes.data,
                                          C.byref(fg), C.byref(ncut)), "gvm_setup_noise_image")
    slices = []
    for f, uvw, vo, w in chans:
        fw.gvm_setup_hermitian_lambda(uvw.shape[0], uvw.ctypes.data, vo.ctypes.data, f)
        slices.append((f, uvw, vo, w))
    f0, u0, v0, w0 = slices[0]
    return Problem(
        M=npix, N=npix, deltax=deltax, deltay=deltay, x0=x0, y0=y0, freq=f0, nu_0=nu0,
        uvw=u0, vo=v0, weight=w0, noise=noise, fg_scale=float(fg.value),
        noise_cut=float(ncut.value), beam_kind=native.GVM_BEAM_GAUSSIAN,
        antenna_diameter=float(np.float32(base.dish_diameter_m)),
        pb_factor=float(np.float32(base.pb_factor)), pb_cutoff=float(np.float32(pb_cutoff)),
        sky=sky, sim=base, slices=slices,
        beam_info=dict(bmaj_deg=bmaj, bmin_deg=bmin, bpa_deg=bpa, sum_weights=sumw,
                       vis_noise=vis_noise, noise_jypix=noise_jypix, flag_fractions=flags),
    )
