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
import json
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
                 H: int = 256, W: int = 256, seed: int = 0, bf16: bool = False):
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
        # BF16 weights: the convolutions read a BF16 copy of the FP32 master
        # parameters (refreshed after every Adam step); biases stay FP32
        self.bf16 = bf16
        self.params_bf16 = DeviceBuffer(2 * off) if bf16 else None
        self._refresh_bf16()
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

    def _wp(self, name):
        """(weight pointer as the convolutions read it, FP32 bias pointer)."""
        wo, bo, *_ = self.offsets[name]
        w = self.params_bf16.ptr + 2 * wo if self.bf16 else self.params.ptr + 4 * wo
        return w, self.params.ptr + 4 * bo

    def _refresh_bf16(self):
        if self.bf16:
            native.check(self.lib.gvm_nn_to_bf16(self.ctx.ctx, self.params.ptr,
                                                 self.params_bf16.ptr, self.nparams), "to_bf16")

    def _conv_fwd(self, *a):
        fn = self.lib.gvm_nn_conv_fwd_bf16 if self.bf16 else self.lib.gvm_nn_conv_fwd
        self._check(fn(self.ctx.ctx, *a), "conv")

    def _conv_bwd(self, *a):
        fn = self.lib.gvm_nn_conv_bwd_bf16 if self.bf16 else self.lib.gvm_nn_conv_bwd
        self._check(fn(self.ctx.ctx, *a), "conv bwd")

    def _check(self, st, what):
        native.check(st, what)

    # ---- forward ----
    def forward(self, nb: Optional[int] = None) -> None:
        """self.x -> self.y for the first `nb` images of the batch (default:
        all of them); the images are independent, so fewer cost less."""
        lib, c = self.lib, self.ctx.ctx
        B = self.B if nb is None else int(nb)
        if not 1 <= B <= self.B:
            raise ValueError(f"forward: {B} images for a batch of {self.B}")
        L, ch = self.levels, self.ch
        cin = 1
        for l in range(L):
            h, w = self.hs[l]
            wa, ba = self._wp(f"e{l}a")
            self._conv_fwd(self.enc_in[l].ptr, wa, ba, self.ea[l].ptr, B, cin, ch[l], h, w, 3, 1)
            wb, bb = self._wp(f"e{l}b")
            self._conv_fwd(self.ea[l].ptr, wb, bb, self.eb[l].ptr, B, ch[l], ch[l], h, w, 3, 1)
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
            wa, ba = self._wp(f"d{l}a")
            self._conv_fwd(self.cat[l].ptr, wa, ba, self.da[l].ptr, B, ccat, ch[l], h, w, 3, 1)
            wb, bb = self._wp(f"d{l}b")
            self._conv_fwd(self.da[l].ptr, wb, bb, self.db[l].ptr, B, ch[l], ch[l], h, w, 3, 1)
            d = self.db[l]
        self._dout = d
        wo, bo = self._wp("out")
        self._conv_fwd(d.ptr, wo, bo, self.y.ptr, B, ch[0], 1, self.H, self.W, 1, 0)

    # ---- backward (after forward; self.g_y holds dL/dy) ----
    def backward(self) -> None:
        lib, c, B = self.lib, self.ctx.ctx, self.B
        L, ch = self.levels, self.ch
        H, W = self.H, self.W
        wo, _ = self._wp("out")
        gwo, gbo = self._wb("out", self.grads)
        self._conv_bwd(self._dout.ptr, wo, None, self.g_y.ptr, self.g_d[0].ptr, gwo, gbo, B, ch[0],
                       1, H, W, 1, 0)
        for l in range(0, L - 1):
            h, w = self.hs[l]
            hs, ws = self.hs[l + 1]
            ccat = ch[l + 1] + ch[l]
            wb, _ = self._wp(f"d{l}b")
            gwb, gbb = self._wb(f"d{l}b", self.grads)
            self._conv_bwd(self.da[l].ptr, wb, self.db[l].ptr, self.g_d[l].ptr, self.g_a[l].ptr,
                           gwb, gbb, B, ch[l], ch[l], h, w, 3, 1)
            wa, _ = self._wp(f"d{l}a")
            gwa, gba = self._wb(f"d{l}a", self.grads)
            self._conv_bwd(self.cat[l].ptr, wa, self.da[l].ptr, self.g_a[l].ptr, self.g_cat[l].ptr,
                           gwa, gba, B, ccat, ch[l], h, w, 3, 1)
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
            wb, _ = self._wp(f"e{l}b")
            gwb, gbb = self._wb(f"e{l}b", self.grads)
            self._conv_bwd(self.ea[l].ptr, wb, self.eb[l].ptr, self.g_e[l].ptr, self.g_a[l].ptr,
                           gwb, gbb, B, ch[l], ch[l], h, w, 3, 1)
            wa, _ = self._wp(f"e{l}a")
            gwa, gba = self._wb(f"e{l}a", self.grads)
            gx = self.g_x[l].ptr if l > 0 else None
            self._conv_bwd(self.enc_in[l].ptr, wa, self.ea[l].ptr, self.g_a[l].ptr, gx, gwa, gba,
                           B, cin, ch[l], h, w, 3, 1)
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
        self._refresh_bf16()

    # ---- host helpers ----
    def set_input(self, xb: np.ndarray) -> None:
        self.x.upload(np.ascontiguousarray(xb, np.float32))

    def output(self, nb: Optional[int] = None) -> np.ndarray:
        """The first `nb` outputs (default: the whole batch), (nb, 1, H, W)."""
        self.ctx.sync()
        nb = self.B if nb is None else int(nb)
        n = nb * self.H * self.W
        return self.y.download(np.float32)[:n].reshape(nb, 1, self.H, self.W)

    def params_dict(self) -> dict:
        """The parameters as the convolutions use them (BF16-rounded weights
        when bf16), in FP64."""
        self.ctx.sync()
        flat = self.params.download(np.float32).astype(np.float64)
        if self.bf16:
            hb = self.params_bf16.download(np.uint16)[: self.nparams]
            wflat = (hb.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        else:
            wflat = flat
        out = {}
        for name, co, ci, k in self.shapes:
            wo, bo, *_ = self.offsets[name]
            out[name] = (wflat[wo:bo].reshape(co, ci, k, k), flat[bo:bo + co])
        return out

    def bias_mask(self) -> np.ndarray:
        """True at the bias entries of the flat parameter vector."""
        m = np.zeros(self.nparams, bool)
        for wo, bo, co, _, _ in self.offsets.values():
            m[bo:bo + co] = True
        return m

    def set_params(self, flat: np.ndarray) -> None:
        """Replaces the FP32 master parameters (flat, storage order) and
        refreshes the BF16 copy the convolutions read."""
        flat = np.ascontiguousarray(flat, np.float32)
        if flat.shape != (self.nparams,):
            raise ValueError(f"set_params: {flat.shape} for {self.nparams} parameters")
        self.params.upload(flat)
        self._refresh_bf16()

    def grads_dict(self) -> dict:
        self.ctx.sync()
        flat = self.grads.download(np.float32).astype(np.float64)
        out = {}
        for name, co, ci, k in self.shapes:
            wo, bo, *_ = self.offsets[name]
            out[name] = (flat[wo:bo].reshape(co, ci, k, k), flat[bo:bo + co])
        return out

    def close(self) -> None:
        for b in [self.params, self.grads, self.m, self.v, self.params_bf16]:
            if b is not None:
                b.free()


# ---------------------------------------------------------------------------
# Training pairs from the plain solver, training, and the warm start itself
# ---------------------------------------------------------------------------

# the objective of the 256^2 workload: chi2 + Entropy (lambda 0.01, G = 0.001)
TERMS = [("Chi2", -1, 0.0), ("Entropy", 0, 0.001)]
LAMBDAS = (0.01,)
# whether the solver replays evaluation graphs (the default of solve())
GRAPHS = True


def observation(seed: int, which: int = 0, **overrides):
    """An observation of workload `which` (0: 256^2 / 36,000 visibilities; 1:
    1024^2 / 1,764,000; 4: 512^2 / 1,029,000) with the sky and the noise drawn
    from `seed` (the array layout stays fixed)."""
    from .problem import make_problem

    kw = dict(seed_sky=seed, seed_noise=seed + 7919)
    kw.update(overrides)
    return make_problem(which, **kw)


def dirty_image(ctx: Context, p) -> np.ndarray:
    """-d chi2 / d I at the default initial image (image 0), float32 (N, M)."""
    d = ctx.image(p.initial_image())
    ctx.chi2(d)
    r = ctx.zeros(2 * p.M * p.N)
    ctx.dchi2(d, r)
    ctx.sync()
    g = r.download(np.float32, (2, p.N, p.M))[0]
    d.free()
    r.free()
    return -g


def solve(p, init: Optional[np.ndarray] = None, iterations: int = 400,
          optimizer: str = "CG-FRPRMN", graphs: Optional[bool] = None, adjoint: str = "bf16x3",
          terms=None, lambdas=None):
    """The plain MEM optimisation (CG-FRPRMN, the workload's optimiser, or
    CG-LBFGS with memory 10), run to its own tolerance or `iterations`, from
    the default image or `init` (image 0), on the objective `terms` /
    `lambdas` (default: TERMS / LAMBDAS).  Returns (final image 0, trajectory
    of objective values with the starting value first, seconds to each entry,
    result)."""
    ctx = Context(p, adjoint=adjoint)
    obj = Objective(ctx, fw_config(TERMS if terms is None else terms,
                                   lambdas=LAMBDAS if lambdas is None else lambdas))
    obj.set_graphs(GRAPHS if graphs is None else graphs)
    img = p.initial_image()
    if init is not None:
        img[0] = np.maximum(init, p.minpix).astype(np.float32)
    d = ctx.image(img)
    f0 = obj.value(d)
    res, traj, secs = obj.optimize_timed(d, optimizer=optimizer, iterations=iterations,
                                         lbfgs_k=10 if optimizer == "CG-LBFGS" else 0)
    out = d.download(np.float32, (2, p.N, p.M))[0]
    obj.close()
    ctx.close()
    return out, [f0] + list(traj), [0.0] + list(secs), res


@dataclass
class Pairs:
    dirty: np.ndarray  # (n, N, M)
    model: np.ndarray  # (n, N, M)
    seeds: list = field(default_factory=list)
    seconds: float = 0.0


def make_pairs(seeds, iterations: int = 400, optimizer: str = "CG-FRPRMN",
               which: int = 0) -> Pairs:
    """Training pairs (dirty image, MEM solution) from the plain solver."""
    t0 = time.perf_counter()
    dirty, model = [], []
    for s in seeds:
        p = observation(s, which)
        ctx = Context(p)
        dirty.append(dirty_image(ctx, p))
        ctx.close()
        model.append(solve(p, iterations=iterations, optimizer=optimizer)[0])
    return Pairs(np.stack(dirty), np.stack(model), list(seeds), time.perf_counter() - t0)


class WarmStart:
    """A trained U-Net and its input/output scaling."""

    def __init__(self, ctx: Context, levels: int = 4, base: int = 16, batch: int = 8,
                 npix: int = 256, seed: int = 0, bf16: bool = True):
        self.net = UNet(ctx, levels=levels, base=base, batch=batch, H=npix, W=npix, seed=seed,
                        bf16=bf16)
        self.ctx = ctx
        self.calib = 1.0
        self.history: list = []
        self.meta: dict = {}

    def _scale(self, dirty):
        return np.maximum(np.abs(dirty).reshape(len(dirty), -1).max(axis=1), 1e-30)

    def train(self, pairs: Pairs, steps: int = 400, lr: float = 1e-3, seed: int = 0,
              log_every: int = 50, progress=None) -> float:
        """Adam on the MSE of the scaled model; returns the seconds taken.
        `progress(step, loss)` is called with every logged loss."""
        n = len(pairs.dirty)
        s = self._scale(pairs.dirty)
        ms = np.maximum(np.abs(pairs.model).reshape(n, -1).max(axis=1), 1e-30)
        self.calib = float(np.median(ms / s))
        x = (pairs.dirty / s[:, None, None]).astype(np.float32)
        y = (pairs.model / (s * self.calib)[:, None, None]).astype(np.float32)
        B, H, W = self.net.B, self.net.H, self.net.W
        target = self.ctx.zeros(B * H * W)
        rng = np.random.default_rng(seed)
        t0 = time.perf_counter()
        for step in range(steps):
            idx = rng.choice(n, B, replace=n < B)
            self.net.set_input(x[idx][:, None])
            target.upload(y[idx][:, None])
            self.net.forward()
            loss = self.net.loss(target)
            self.net.backward()
            self.net.adam(lr=lr)
            if step % log_every == 0 or step == steps - 1:
                self.history.append((step, loss))
                if progress:
                    progress(step, loss)
        self.ctx.sync()
        target.free()
        return time.perf_counter() - t0

    def predict(self, dirty: np.ndarray) -> np.ndarray:
        """Initial model images for a stack of dirty images (n, N, M)."""
        B = self.net.B
        out = []
        s = self._scale(dirty)
        x = (dirty / s[:, None, None]).astype(np.float32)
        for k in range(0, len(x), B):
            xk = x[k:k + B]
            self.net.set_input(xk[:, None])
            self.net.forward(len(xk))
            out.extend(self.net.output(len(xk))[:, 0])
        return np.stack(out) * (s * self.calib)[:, None, None]

    def save(self, path, meta: Optional[dict] = None) -> None:
        """Writes the network as its convolutions use it to an .npz file that
        np.load reads without pickle: the BF16 weight bits and the FP32
        biases (every parameter in FP32 for a network with FP32 weights), the
        shape, the calibration constant and `meta` (a JSON object: how the
        network was trained)."""
        net = self.net
        self.ctx.sync()
        flat = net.params.download(np.float32)[: net.nparams]
        bias = net.bias_mask()
        arrays = dict(levels=np.int64(net.levels), base=np.int64(net.base),
                      npix=np.int64(net.H), calib=np.float64(self.calib), bf16=np.bool_(net.bf16),
                      meta=np.str_(json.dumps(meta or {})))
        if net.bf16:
            bits = net.params_bf16.download(np.uint16)[: net.nparams]
            arrays["weights_bf16"] = bits[~bias]
            arrays["biases"] = flat[bias]
        else:
            arrays["params"] = flat
        np.savez_compressed(path, **arrays)

    @classmethod
    def load(cls, ctx: Context, path, batch: int = 1, npix: Optional[int] = None) -> "WarmStart":
        """A network written by save(), for batches of `batch` images of
        npix x npix (default: the size it was trained at).  Its predictions
        equal those of the network that was saved, bit for bit."""
        with np.load(path, allow_pickle=False) as z:
            levels, base = int(z["levels"]), int(z["base"])
            ws = cls(ctx, levels=levels, base=base, batch=batch,
                     npix=int(z["npix"]) if npix is None else int(npix), bf16=bool(z["bf16"]))
            net = ws.net
            if net.bf16:
                bias = net.bias_mask()
                flat = np.empty(net.nparams, np.float32)
                flat[~bias] = (z["weights_bf16"].astype(np.uint32) << 16).view(np.float32)
                flat[bias] = z["biases"]
            else:
                flat = z["params"]
            net.set_params(flat)
            ws.calib = float(z["calib"])
            ws.meta = json.loads(str(z["meta"]))
        return ws

    def close(self) -> None:
        self.net.close()


def start_image(ws: WarmStart, ctx: Context, obj: Objective, p):
    """Where the MEM optimisation starts: the network's prediction from the
    dirty image (clipped to minpix, alpha 0) when the objective `obj` is
    lower there than at the default image, else the default image, the
    fallback.  Returns (device image, True when the prediction is used,
    objective value at the start)."""
    dirty = dirty_image(ctx, p)
    img = p.initial_image()
    img[0] = np.maximum(ws.predict(dirty[None])[0], p.minpix).astype(np.float32)
    d = ctx.image(img)
    f_pred = obj.value(d)
    d0 = ctx.image(p.initial_image())
    f_default = obj.value(d0)
    if f_pred < f_default:
        d0.free()
        return d, True, f_pred
    d.free()
    return d0, False, f_default


def time_to(traj, secs, target: float):
    """(iterations, seconds) to the first trajectory entry <= target (entry 0
    is the start), or None."""
    for k, (f, t) in enumerate(zip(traj, secs)):
        if f <= target:
            return k, t
    return None


def evaluate(ws: WarmStart, seeds, iterations: int = 400, optimizer: str = "CG-FRPRMN",
             plain: bool = False, which: int = 0, progress=None, terms=None,
             lambdas=None) -> dict:
    """The plain solver from the default image and from the warm start on
    held-out observations: iterations and seconds (dirty image and inference
    included for the warm start) to the objective value both reach, and the
    relative L2 error of both final images against the true sky inside the
    mask.  With plain=True, also the plain FP32 form (FP32 adjoint, no
    graphs, default image) to the value all three reach."""
    rows = []
    for s in seeds:
        p = observation(s, which)
        ctx = Context(p)
        t0 = time.perf_counter()
        dirty = dirty_image(ctx, p)
        init = ws.predict(dirty[None])[0]
        t_pre = time.perf_counter() - t0
        ctx.close()
        obj_kw = dict(terms=terms, lambdas=lambdas)
        img_d, tr_d, sc_d, _ = solve(p, iterations=iterations, optimizer=optimizer, **obj_kw)
        img_w, tr_w, sc_w, _ = solve(p, init=init, iterations=iterations, optimizer=optimizer,
                                     **obj_kw)
        # final images against the true sky (internal units) inside the mask
        mask = p.noise < p.noise_cut
        truth = (p.sky / p.fg_scale)[mask]
        err_d = float(np.linalg.norm(img_d[mask] - truth) / np.linalg.norm(truth))
        err_w = float(np.linalg.norm(img_w[mask] - truth) / np.linalg.norm(truth))
        diff = float(np.linalg.norm(img_w[mask] - img_d[mask]) / np.linalg.norm(img_d[mask]))
        target = max(tr_d[-1], tr_w[-1])
        extra = {}
        if plain:
            img_p, tr_p, sc_p, _ = solve(p, iterations=iterations, optimizer=optimizer,
                                         graphs=False, adjoint="f32", **obj_kw)
            t3 = max(target, tr_p[-1])
            cp, cw = time_to(tr_p, sc_p, t3), time_to(tr_w, sc_w, t3)
            extra = dict(f_plain=tr_p[-1], iters_plain=cp[0], seconds_plain=cp[1],
                         seconds_warm_to_common=cw[1] + t_pre,
                         sky_err_plain=float(np.linalg.norm(img_p[mask] - truth) /
                                             np.linalg.norm(truth)))
        a, b = time_to(tr_d, sc_d, target), time_to(tr_w, sc_w, target)
        rows.append(dict(seed=s, f_default=tr_d[-1], f_warm=tr_w[-1], f0_default=tr_d[0],
                         f0_warm=tr_w[0],
                         iters_default=a[0], seconds_default=a[1],
                         iters_warm=b[0], seconds_warm=b[1] + t_pre, warm_overhead_s=t_pre,
                         iters_run_default=len(tr_d), iters_run_warm=len(tr_w),
                         sky_err_default=err_d, sky_err_warm=err_w, image_rel_diff=diff, **extra))
        if progress:
            progress(rows[-1])
    it_d = np.array([r["iters_default"] for r in rows], float)
    it_w = np.maximum(np.array([r["iters_warm"] for r in rows], float), 1.0)
    t_d = np.array([r["seconds_default"] for r in rows])
    t_w = np.array([r["seconds_warm"] for r in rows])
    out = {}
    if plain:
        out["median_sky_err_plain"] = float(np.median([r["sky_err_plain"] for r in rows]))
        out["median_time_ratio_plain_fp32"] = float(np.median(
            [r["seconds_plain"] / r["seconds_warm_to_common"] for r in rows]))
    out["median_sky_err_default"] = float(np.median([r["sky_err_default"] for r in rows]))
    out["median_sky_err_warm"] = float(np.median([r["sky_err_warm"] for r in rows]))
    out["median_image_rel_diff"] = float(np.median([r["image_rel_diff"] for r in rows]))
    return dict(rows=rows, **out, median_iter_ratio=float(np.median(it_d / it_w)),
                median_time_ratio=float(np.median(t_d / t_w)),
                mean_iters_default=float(it_d.mean()), mean_iters_warm=float(it_w.mean()),
                warm_better_or_equal_final=int(sum(r["f_warm"] <= r["f_default"] * (1 + 1e-6)
                                                   for r in rows)))
