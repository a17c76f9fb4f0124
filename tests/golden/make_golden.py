"""Writes the golden vectors in this folder.

The reference ships no golden vectors or known-answer tests for this path
(its tests check exit codes on measurement sets).  These vectors come from an
independent numpy restatement of the forward model and the adjoint, built
on library and brute-force arithmetic, not on the C oracle:

* the unnormalised inverse 2-D FFT (cuFFT CUFFT_INVERSE, src/functions.cu:
  2153-2178) is ``numpy.fft.ifft2(x) * (N * M)``;
* phase rotation (src/functions.cu:2483-2518), bilinear degridding with the
  off-grid rule (src/functions.cu:2557-2609), residual and chi2
  (src/functions.cu:2663-2671, 2867-2876, 4323-4454) are written out per
  visibility in FP64;
* the gradient DChi2 (src/functions.cu:3698-3791) and its chain rule
  (src/functions.cu:3968-4024) are a brute-force FP64 direct sum over every
  pixel and visibility;
* the Gaussian primary beam (src/functions.cu:2293-2333) and clip2IWNoise /
  calculateInu (src/functions.cu:2694-2719, 3939-3966) by their formulas.

Run ``python tests/golden/make_golden.py`` to regenerate ``forward_adjoint_64.npz``.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIGHTSPEED = 2.99792458e8


def geometry():
    npix = 64
    return dict(
        M=npix, N=npix, deltax=-0.05 / 3600.0, deltay=0.05 / 3600.0, x0=32.0, y0=32.0,
        freq=float(np.float32(230e9)), nu_0=float(np.float32(230e9 / 1.05)),
        beam_kind=1, antenna_diameter=12.0, pb_factor=1.13, pb_cutoff=10.0,
        minpix=0.001, eta=-1.0, threshold=0.0, fg_scale=0.375, noise_cut=2.0,
    )


def primary_beam(g):
    n = g["N"]
    lam = LIGHTSPEED / g["freq"]
    fwhm = g["pb_factor"] * lam / g["antenna_diameter"]
    j = np.arange(n)
    x = (j - np.floor(g["x0"])) * g["deltax"] * np.pi / 180.0
    y = (j - np.floor(g["y0"])) * g["deltay"] * np.pi / 180.0
    r = np.sqrt(x[None, :] ** 2 + y[:, None] ** 2)
    pb = np.exp(-4.0 * np.log(2.0) * (r / fwhm) ** 2)
    pb[r > g["pb_cutoff"]] = 0.0
    return pb


def clip_and_inu(img, noise, g):
    """clip2IWNoise then calculateInu on one channel: returns (clipped image, I_nu).
    MINPIX is a float in the reference, so the floor is -eta * float(MINPIX)."""
    floor = -g["eta"] * float(np.float32(g["minpix"]))
    out = img.copy()
    masked = noise > g["noise_cut"]
    out[0][masked] = floor
    out[1][masked] = 0.0
    inu = out[0] * (g["freq"] / g["nu_0"]) ** out[1]
    inu = np.where(inu < floor, floor, inu)
    return out, inu


def model_grid(inu, pb, g):
    n, m = g["N"], g["M"]
    v = np.fft.ifft2(inu * pb * g["fg_scale"]) * (n * m)
    js = np.arange(m)
    js = np.where(js < m // 2, js, js - m)
    i_s = np.arange(n)
    i_s = np.where(i_s < n // 2, i_s, i_s - n)
    u = g["x0"] / m * js
    vv = g["y0"] / n * i_s
    return v * np.exp(-2j * np.pi * (u[None, :] + vv[:, None]))


def degrid(grid, uvw, weight, g):
    n = g["N"]
    du = 1.0 / (g["M"] * g["deltax"] * np.pi / 180.0)
    dv = 1.0 / (g["N"] * g["deltay"] * np.pi / 180.0)
    vm = np.zeros(uvw.shape[0], np.complex128)
    w = weight.astype(np.float64).copy()
    for k in range(uvw.shape[0]):
        ux = uvw[k, 0] / du
        uy = uvw[k, 1] / dv
        if ux < 0:
            ux += n
        if uy < 0:
            uy += n
        i1, j1 = int(np.floor(ux)), int(np.floor(uy))
        if not (0 <= i1 < n and 0 <= j1 < n):
            w[k] = 0.0
            continue
        i2, j2 = (i1 + 1) % n, (j1 + 1) % n
        a, b = ux - i1, uy - j1
        vm[k] = ((1 - a) * (1 - b) * grid[j1, i1] + a * (1 - b) * grid[j1, i2]
                 + (1 - a) * b * grid[j2, i1] + a * b * grid[j2, i2])
    return vm, w


def dchi2_direct(vr, uvw, w, clipped, noise, pb, g, flag_opt):
    n, m = g["N"], g["M"]
    j = np.arange(m)
    i = np.arange(n)
    x = (j - np.floor(g["x0"])) * g["deltax"] * np.pi / 180.0
    y = (i - np.floor(g["y0"])) * g["deltay"] * np.pi / 180.0
    X, Y = np.meshgrid(x, y)
    Z = np.sqrt(1.0 - X ** 2 - Y ** 2)
    phase = 2.0 * np.pi * (X[..., None] * uvw[:, 0] + Y[..., None] * uvw[:, 1]
                           + (Z[..., None] - 1.0) * uvw[:, 2])
    s = np.sum(w * (vr.real * np.cos(phase) + vr.imag * np.sin(phase)), axis=-1)
    d = -s * g["fg_scale"] * pb
    d[~(noise < g["noise_cut"])] = 0.0
    nudiv = g["freq"] / g["nu_0"]
    res = np.zeros((2, n, m))
    i0, alpha = clipped
    if flag_opt % 2 == 0:
        res[0] = d * nudiv ** alpha
    else:
        res[1] = np.where(i0 > g["threshold"],
                          d * i0 * nudiv ** alpha * g["fg_scale"] * np.log(nudiv), 0.0)
    res[:, ~(noise < g["noise_cut"])] = 0.0
    return res


def make():
    g = geometry()
    rng = np.random.default_rng(2024)
    n = g["N"]
    nvis = 300
    du = 1.0 / (g["M"] * abs(g["deltax"]) * np.pi / 180.0)
    uvw = np.zeros((nvis, 3))
    uvw[:, 0] = -rng.uniform(0.0, 31.0, nvis) * du          # Hermitian half: u <= 0
    uvw[:, 1] = rng.uniform(-31.0, 31.0, nvis) * du
    uvw[:, 2] = rng.uniform(-3e4, 3e4, nvis)
    uvw[::37, 0] = -70.3 * du                                 # column 70.3 >= N: off the grid
    vo = (rng.standard_normal((nvis, 2)) * 0.05).astype(np.float32)
    weight = rng.uniform(0.5, 2.0, nvis).astype(np.float32)
    noise = np.full((n, n), 1.0, np.float32)
    noise[:3, :] = 5.0                                        # outside the mask
    noise[:, -2:] = 5.0
    img = np.zeros((2, n, n), np.float32)
    img[0] = rng.uniform(0.0, 2.0, (n, n)).astype(np.float32)
    img[0][rng.random((n, n)) < 0.1] = 1e-4                   # below the floor: clipped by I_nu
    img[1] = rng.uniform(-1.0, 2.0, (n, n)).astype(np.float32)

    pb = primary_beam(g)
    clipped, inu = clip_and_inu(img.astype(np.float64), noise, g)
    grid = model_grid(inu, pb, g)
    vm, w = degrid(grid, uvw, weight, g)
    vr = vo[:, 0] + 1j * vo[:, 1] - vm
    chi2 = 0.5 * np.sum(w * np.abs(vr) ** 2)
    g0 = dchi2_direct(vr, uvw, w, clipped, noise, pb, g, 0)
    g1 = dchi2_direct(vr, uvw, w, clipped, noise, pb, g, 1)
    np.savez_compressed(
        HERE / "forward_adjoint_64.npz",
        geometry=np.array([repr(g)]), uvw=uvw, vo=vo, weight=weight, noise=noise, image=img,
        pb=pb, clipped=clipped, vr=vr, chi2=np.array(chi2), weight_after=w,
        grad_flag0=g0, grad_flag1=g1,
    )
    # the unnormalised inverse FFT on a non-square grid
    x = rng.standard_normal((8, 16)) + 1j * rng.standard_normal((8, 16))
    np.savez_compressed(HERE / "fft_inverse_8x16.npz", x=x, y=np.fft.ifft2(x) * x.size)


if __name__ == "__main__":
    make()
