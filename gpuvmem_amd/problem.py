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

    @property
    def nvis(self) -> int:
        return int(self.uvw.shape[0])

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
    1,764,000 vis) with optional field overrides and runs the set-up."""
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
