"""Gridded-visibility mode: gridding kernels, the device gridder and the
weighting schemes (reference MFS::configure, src/mfs.cu:499-527).

The computation runs in libgvm_hip (``gvm_grid_visibilities_dev``,
``gvm_weighting_apply_dev``, ``gvm_set_gcf``); this module builds the
arguments.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import native


@dataclass
class GriddingKernel:
    """A CKernel's taps (m rows x n columns, float32) and supports.

    The reference derives both supports from m (CKernel::setSupports,
    include/classes/ckernel.cuh:508-511)."""

    taps: np.ndarray
    support_x: int
    support_y: int
    name: str = ""
    gcf: Optional[np.ndarray] = None  # (N, M) float32 correction image, None: all ones

    @classmethod
    def from_taps(cls, taps: np.ndarray, name: str = "", gcf=None) -> "GriddingKernel":
        taps = np.ascontiguousarray(taps, np.float32)
        s = int(np.floor(taps.shape[0] / 2.0))
        return cls(taps, s, s, name, gcf)

    @classmethod
    def pillbox(cls) -> "GriddingKernel":
        """PillBox2D() (src/pillBox2D.cu:22-25): 1 x 1, tap 1, GCF 1."""
        return cls.from_taps(np.ones((1, 1), np.float32), "Pill Box")

    @classmethod
    def build(cls, kind: str, M: int, N: int, deltax: float, deltay: float,
              with_gcf: bool = True, **params) -> "GriddingKernel":
        """A kernel of the reference's CKernel family built as MFS::configure
        does (src/mfs.cu:510-516): taps with sigmas (|deltau|, |deltav|), and
        the correction image for an M x N image of cells (deltax, deltay)
        degrees.  params override the default constructor's m, n, amp, w,
        w2, alpha.  Sinc2D and GaussianSinc2D have no correction image."""
        fw = native.load_fw()
        p = native.CKernelParams()
        if fw.gvm_ckernel_defaults(native.CKERNEL_KINDS[kind], C.byref(p)) != 0:
            raise ValueError(f"unknown kernel {kind}")
        for k, v in params.items():
            setattr(p, k, v)
        dxr, dyr = deltax * np.pi / 180.0, deltay * np.pi / 180.0
        du, dv = 1.0 / (M * dxr), 1.0 / (N * dyr)
        taps = np.zeros((p.m, p.n), np.float32)
        sx, sy = C.c_int(), C.c_int()
        if fw.gvm_ckernel_build(C.byref(p), abs(du), abs(dv), taps.ctypes.data, C.byref(sx),
                                C.byref(sy)) != 0:
            raise ValueError("gvm_ckernel_build failed")
        gcf = None
        if with_gcf:
            gcf = np.zeros((M, N), np.float32)
            if fw.gvm_ckernel_gcf(C.byref(p), M, N, abs(dxr), abs(dyr), gcf.ctypes.data) != 0:
                raise ValueError(f"the {kind} kernel has no correction image")
        return cls(taps, sx.value, sy.value, kind, gcf)

    def as_c(self) -> native.CKernel:
        k = native.CKernel()
        k.m, k.n = self.taps.shape
        k.support_x, k.support_y = self.support_x, self.support_y
        k.taps = self.taps.ctypes.data
        return k


def grid_visibilities_dev(ctx, kernel: GriddingKernel, freq: float, uvw_m, vo, weight, nvis: int):
    """Grids one slice of raw rows held in device buffers (metres, before the
    Hermitian flip).  Returns (uvw (n,3) float64 metres, vo (n,2) float32,
    weight (n,) float32) copied to the host."""
    from .device import DeviceBuffer

    cap = ctx.n
    ou, ov, ow = DeviceBuffer(24 * cap), DeviceBuffer(8 * cap), DeviceBuffer(4 * cap)
    nout = C.c_long()
    kc = kernel.as_c()
    native.check(ctx.lib.gvm_grid_visibilities_dev(ctx.ctx, C.byref(kc), freq, nvis, uvw_m.ptr,
                                                   vo.ptr, weight.ptr, ou.ptr, ov.ptr, ow.ptr, cap,
                                                   C.byref(nout)), "gvm_grid_visibilities_dev")
    n = nout.value
    u = ou.download(np.float64)[: 3 * n].reshape(n, 3).copy()
    v = ov.download(np.float32)[: 2 * n].reshape(n, 2).copy()
    w = ow.download(np.float32)[:n].copy()
    for b in (ou, ov, ow):
        b.free()
    return u, v, w


def add_gridded_dev(ctx, kernel: GriddingKernel, freq: float, uvw_m, vo, weight,
                    nvis: int) -> tuple[int, int]:
    """Grids one slice on the device and adds the gridded rows to ctx as a
    slice.  Returns (slice index, gridded row count)."""
    s = C.c_int()
    ng = C.c_long()
    kc = kernel.as_c()
    native.check(ctx.lib.gvm_add_visibilities_gridded_dev(ctx.ctx, C.byref(kc), freq, nvis,
                                                          uvw_m.ptr, vo.ptr, weight.ptr,
                                                          C.byref(ng), C.byref(s)),
                 "gvm_add_visibilities_gridded_dev")
    return s.value, ng.value


def apply_weighting_dev(ctx, scheme: str, slices: Sequence, robust: float = 0.0,
                        taper: Optional[dict] = None) -> None:
    """WeightingScheme::apply on device rows: slices = [(freq, uvw_m buffer,
    weight buffer, nvis)]; the weight buffers are updated in place."""
    nsl = len(slices)
    freqs = (C.c_float * nsl)(*[s[0] for s in slices])
    nv = (C.c_long * nsl)(*[s[3] for s in slices])
    up = (C.c_void_p * nsl)(*[s[1].ptr for s in slices])
    wp = (C.c_void_p * nsl)(*[s[2].ptr for s in slices])
    t = native.UVTaper(**taper) if taper is not None else None
    native.check(ctx.lib.gvm_weighting_apply_dev(ctx.ctx, native.WEIGHTING_SCHEMES[scheme], robust,
                                                 C.byref(t) if t is not None else None, nsl, freqs,
                                                 nv, up, wp), "gvm_weighting_apply_dev")


def set_gcf(ctx, gcf: Optional[np.ndarray]) -> None:
    """Sets (or clears, with None) the gridding correction image of ctx."""
    if gcf is None:
        native.check(ctx.lib.gvm_set_gcf(ctx.ctx, None), "gvm_set_gcf")
        return
    g = np.ascontiguousarray(gcf, np.float32)
    if g.size != ctx.n:
        raise ValueError("gcf must hold M*N values")
    native.check(ctx.lib.gvm_set_gcf(ctx.ctx, g.ctypes.data), "gvm_set_gcf")


def configure_gridded(ctx, kernel: GriddingKernel, slices: Sequence, scheme: str = "natural",
                      robust: float = 0.0, taper: Optional[dict] = None) -> list:
    """The gridded branch of MFS::configure (src/mfs.cu:499-527) on device
    rows: the weighting scheme over all slices, then each slice gridded with
    `kernel` and added to ctx, then the kernel's correction image.  slices =
    [(freq, uvw_m buffer, vo buffer, weight buffer, nvis)] (raw rows in
    metres; the weight buffers are updated in place).  Returns the gridded
    row count of each slice."""
    apply_weighting_dev(ctx, scheme, [(f, u, w, n) for f, u, _, w, n in slices], robust, taper)
    counts = []
    for f, u, v, w, n in slices:
        counts.append(add_gridded_dev(ctx, kernel, f, u, v, w, n)[1])
    set_gcf(ctx, kernel.gcf)
    return counts
