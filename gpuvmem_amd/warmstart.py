"""Learned warm start: a small U-Net that maps the dirty image to an initial
model for the MEM optimiser, trained on pairs the plain solver produces.

No counterpart in the reference, which starts every optimisation from a
constant image (src/mfs.cu:737-745).  Every layer runs on the device kernels
of include/gvm_nn.h (forward, backward, loss, Adam); this module holds the
network's buffers and the order of the calls.  The optimisation that
follows is the plain MEM optimisation, unchanged: the network only chooses
where it starts, and the default constant image stays the fallback.

Network: `levels` levels of two 3x3 convolutions with ReLU (channels base,
2 base, ...), 2x2 average pooling down, nearest 2x upsampling up with the
skip connection concatenated, and a 1x1 output convolution.  Input: the
dirty image (the negative chi2 gradient at the initial image) divided by its
maximum; output: the model image over the same maximum times a fixed
calibration constant (learned from the training set).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import native
from .device import Context, DeviceBuffer, Objective, fw_config


def unet_shapes(levels: int, base: int, cin: int = 1, cout: int = 1):
    """Parameters in storage order: (name, Co, Ci, K)."""
    ch = [base * 2 ** l for l in range(levels)]
    shapes = []
    c_in = cin
    for l in range(levels):
        shapes.append((f"e{l}a", ch[l], c_in, 3))
        shapes.append((f"e{l}b", ch[l], ch[l], 3))
        c_in = ch[l]
    for l in range(levels - 2, -1, -1):
        shapes.append((f"d{l}a", ch[l], ch[l + 1] + ch[l], 3))
        shapes.append((f"d{l}b", ch[l], ch[l], 3))
    shapes.append(("out", cout, ch[0], 1))
    return shapes


class UNet:
    """Device buffers and call order of the U-Net for batches of B images of
    H x W (H, W divisible by 2^(levels-1))."""

    def __init__(self, ctx: Context, levels: int = 4, base: int = 16, batch: int = 8,
                 H: int = 256, W: int = 256, seed: int = 0):
        if H % (1 << (levels - 1)) or W % (1 << (levels - 1)):
            raise ValueError("image size must be divisible by 2^(levels-1)")
        self.ctx, self.lib = ctx, ctx.lib
        self.levels, self.base, self.B, self.H, self.W = levels, base, batch, H, W
        self.shapes = unet_shapes(levels, base)
        # parameter offsets in one flat buffer: w then b per layer
        self.offsets = {}
        off = 0
        for name, co, ci, k in self.shapes:
            self.offsets[name] = (off, off + co * ci * k * k, co, ci, k)
            off += co * ci * k * k + co
        self.nparams = off
        rng = np.random.default_rng(seed)
        flat = np.zeros(off, np.float32)
        for name, co, ci, k in self.shapes:
            wo, bo, *_ = self.offsets[name]
            std = np.sqrt(2.0 / (ci * k * k)) if name != "out" else np.sqrt(1.0 / (ci * k * k))
            flat[wo:bo] = (rng.standard_normal(co * ci * k * k) * std).astype(np.float32)
        self.params = DeviceBuffer.from_array(flat)
        self.grads = ctx.zeros(off)
        self.m = ctx.zeros(off)
        self.v = ctx.zeros(off)
        self.step = 0
        self._alloc()

    # ---- buffers ----
    def _buf(self, c, h, w):
        return self.ctx.zeros(self.B * c * h * w)

    def _alloc(self):
        L, B = self.levels, self.B
        ch = [self.base * 2 ** l for l in range(L)]
        self.ch = ch
        self.hs = [(self.H >> l, self.W >> l) for l in range(L)]
        self.x = self._buf(1, self.H, self.W)
        self.enc_in = [self.x] + [self._buf(ch[l - 1], *self.hs[l]) for l in range(1, L)]
        self.ea = [self._buf(ch[l], *self.hs[l]) for l in range(L)]
        self.eb = [self._buf(ch[l], *self.hs[l]) for l in range(L)]
        self.cat = {l: self._buf(ch[l + 1] + ch[l], *self.hs[l]) for l in range(L - 1)}
        self.da = {l: self._buf(ch[l], *self.hs[l]) for l in range(L - 1)}
        self.db = {l: self._buf(ch[l], *self.hs[l]) for l in range(L - 1)}
        self.y = self._buf(1, self.H, self.W)
        # gradients
        self.g_y = self._buf(1, self.H, self.W)
        self.g_d = {l: self._buf(ch[l], *self.hs[l]) for l in range(L)}  # at decoder outputs
        self.g_a = [self._buf(ch[l], *self.hs[l]) for l in range(L)]
        self.g_cat = {l: self._buf(ch[l + 1] + ch[l], *self.hs[l]) for l in range(L - 1)}
        self.g_e = [self._buf(ch[l], *self.hs[l]) for l in range(L)]
        self.g_x = [None] + [self._buf(ch[l - 1], *self.hs[l]) for l in range(1, L)]
        self.g_pool = [self._buf(ch[l], *self.hs[l]) for l in range(L)]

    def _wb(self, name, buf):
        wo, bo, *_ = self.offsets[name]
        return buf.ptr + 4 * wo, buf.ptr + 4 * bo

    def _check(self, st, what):
        native.check(st, what)

    # ---- forward ----
    def forward(self) -> None:
        """self.x -> self.y for the whole batch."""
        lib, c, B = self.lib, self.ctx.ctx, self.B
        L, ch = self.levels, self.ch
        cin = 1
        for l in range(L):
            h, w = self.hs[l]
            wa, ba = self._wb(f"e{l}a", self.params)
            self._check(lib.gvm_nn_conv_fwd(c, self.enc_in[l].ptr, wa, ba, self.ea[l].ptr, B, cin,
                                            ch[l], h, w, 3, 1), "conv")
            wb, bb = self._wb(f"e{l}b", self.params)
            self._check(lib.gvm_nn_conv_fwd(c, self.ea[l].ptr, wb, bb, self.eb[l].ptr, B, ch[l],
                                            ch[l], h, w, 3, 1), "conv")
            if l < L - 1:
                self._check(lib.gvm_nn_pool_fwd(c, self.eb[l].ptr, self.enc_in[l + 1].ptr, B,
                                                ch[l], h, w), "pool")
            cin = ch[l]
        d = self.eb[L - 1]
        for l in range(L - 2, -1, -1):
            h, w = self.hs[l]
            hs, ws = self.hs[l + 1]
            ccat = ch[l + 1] + ch[l]
            self._check(lib.gvm_nn_up_fwd(c, d.ptr, self.cat[l].ptr, B, ch[l + 1], ccat, 0, hs, ws),
                        "up")
            self._check(lib.gvm_nn_copy_channels(c, self.eb[l].ptr, self.cat[l].ptr, B, ch[l],
                                                 ch[l], 0, ccat, ch[l + 1], h * w), "copy")
            wa, ba = self._wb(f"d{l}a", self.params)
            self._check(lib.gvm_nn_conv_fwd(c, self.cat[l].ptr, wa, ba, self.da[l].ptr, B, ccat,
                                            ch[l], h, w, 3, 1), "conv")
            wb, bb = self._wb(f"d{l}b", self.params)
            self._check(lib.gvm_nn_conv_fwd(c, self.da[l].ptr, wb, bb, self.db[l].ptr, B, ch[l],
                                            ch[l], h, w, 3, 1), "conv")
            d = self.db[l]
        self._dout = d
        wo, bo = self._wb("out", self.params)
        self._check(lib.gvm_nn_conv_fwd(c, d.ptr, wo, bo, self.y.ptr, B, ch[0], 1, self.H, self.W,
                                        1, 0), "conv")

    # ---- backward (after forward; self.g_y holds dL/dy) ----
    def backward(self) -> None:
        lib, c, B = self.lib, self.ctx.ctx, self.B
        L, ch = self.levels, self.ch
        H, W = self.H, self.W
        wo, _ = self._wb("out", self.params)
        gwo, gbo = self._wb("out", self.grads)
        self._check(lib.gvm_nn_conv_bwd(c, self._dout.ptr, wo, None, self.g_y.ptr, self.g_d[0].ptr,
                                        gwo, gbo, B, ch[0], 1, H, W, 1, 0), "conv bwd")
        for l in range(0, L - 1):
            h, w = self.hs[l]
            hs, ws = self.hs[l + 1]
            ccat = ch[l + 1] + ch[l]
            wb, _ = self._wb(f"d{l}b", self.params)
            gwb, gbb = self._wb(f"d{l}b", self.grads)
            self._check(lib.gvm_nn_conv_bwd(c, self.da[l].ptr, wb, self.db[l].ptr, self.g_d[l].ptr,
                                            self.g_a[l].ptr, gwb, gbb, B, ch[l], ch[l], h, w, 3, 1),
                        "conv bwd")
            wa, _ = self._wb(f"d{l}a", self.params)
            gwa, gba = self._wb(f"d{l}a", self.grads)
            self._check(lib.gvm_nn_conv_bwd(c, self.cat[l].ptr, wa, self.da[l].ptr, self.g_a[l].ptr,
                                            self.g_cat[l].ptr, gwa, gba, B, ccat, ch[l], h, w, 3, 1),
                        "conv bwd")
            # skip part -> gradient at eb[l]; up part -> gradient at the level below's output
            self._check(lib.gvm_nn_copy_channels(c, self.g_cat[l].ptr, self.g_e[l].ptr, B, ch[l],
                                                 ccat, ch[l + 1], ch[l], 0, h * w), "copy")
            target = self.g_d[l + 1] if l + 1 < L - 1 else self.g_e[L - 1]
            self._check(lib.gvm_nn_up_bwd(c, self.g_cat[l].ptr, target.ptr, B, ch[l + 1], ccat, 0,
                                          hs, ws), "up bwd")
        for l in range(L - 1, -1, -1):
            h, w = self.hs[l]
            cin = ch[l - 1] if l > 0 else 1
            if l < L - 1:
                # eb[l] feeds the skip (g_e[l]) and the pool (g_pool[l])
                self._check(lib.gvm_nn_add(c, self.g_pool[l].ptr, self.g_e[l].ptr,
                                           B * ch[l] * h * w), "add")
            wb, _ = self._wb(f"e{l}b", self.params)
            gwb, gbb = self._wb(f"e{l}b", self.grads)
            self._check(lib.gvm_nn_conv_bwd(c, self.ea[l].ptr, wb, self.eb[l].ptr, self.g_e[l].ptr,
                                            self.g_a[l].ptr, gwb, gbb, B, ch[l], ch[l], h, w, 3, 1),
                        "conv bwd")
            wa, _ = self._wb(f"e{l}a", self.params)
            gwa, gba = self._wb(f"e{l}a", self.grads)
            gx = self.g_x[l].ptr if l > 0 else None
            self._check(lib.gvm_nn_conv_bwd(c, self.enc_in[l].ptr, wa, self.ea[l].ptr,
                                            self.g_a[l].ptr, gx, gwa, gba, B, cin, ch[l], h, w, 3,
                                            1), "conv bwd")
            if l > 0:
                hp, wp = self.hs[l - 1]
                self._check(lib.gvm_nn_pool_bwd(c, self.g_x[l].ptr, self.g_pool[l - 1].ptr, B,
                                                ch[l - 1], hp, wp), "pool bwd")

    def loss(self, target: DeviceBuffer) -> float:
        """MSE of self.y against target; writes self.g_y.  Synchronises."""
        n = self.B * self.H * self.W
        self._check(self.lib.gvm_nn_mse(self.ctx.ctx, self.y.ptr, target.ptr, self.g_y.ptr, n, 0),
                    "mse")
        out = (C.c_double * 1)()
        self._check(self.lib.gvm_values_read(self.ctx.ctx, 1, out), "values")
        return out[0]

    def adam(self, lr: float = 1e-3, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self.step += 1
        self._check(self.lib.gvm_nn_adam(self.ctx.ctx, self.params.ptr, self.grads.ptr, self.m.ptr,
                                         self.v.ptr, self.nparams, lr, b1, b2, eps, self.step),
                    "adam")

    # ---- host helpers ----
    def set_input(self, xb: np.ndarray) -> None:
        self.x.upload(np.ascontiguousarray(xb, np.float32))

    def output(self) -> np.ndarray:
        self.ctx.sync()
        return self.y.download(np.float32, (self.B, 1, self.H, self.W))

    def params_dict(self) -> dict:
        flat = self.params.download(np.float32).astype(np.float64)
        out = {}
        for name, co, ci, k in self.shapes:
            wo, bo, *_ = self.offsets[name]
            out[name] = (flat[wo:bo].reshape(co, ci, k, k), flat[bo:bo + co])
        return out

    def grads_dict(self) -> dict:
        self.ctx.sync()
        flat = self.grads.download(np.float32).astype(np.float64)
        out = {}
        for name, co, ci, k in self.shapes:
            wo, bo, *_ = self.offsets[name]
            out[name] = (flat[wo:bo].reshape(co, ci, k, k), flat[bo:bo + co])
        return out

    def close(self) -> None:
        for b in [self.params, self.grads, self.m, self.v]:
            b.free()
