"""FP64 numpy restatement of the warm-start U-Net's layers and of the whole
network's forward and backward pass.

TEST INFRASTRUCTURE ONLY: used by tests/ to check the device kernels of
include/gvm_nn.h and the network built on them (gpuvmem_amd/warmstart.py).
The layers have no counterpart in the reference.  Layouts are NCHW.
"""
from __future__ import annotations

import numpy as np


def conv_fwd(x, w, b, relu):
    """x [B,Ci,H,W], w [Co,Ci,K,K], b [Co]; zero padding K//2."""
    B, Ci, H, W = x.shape
    Co, _, K, _ = w.shape
    P = K // 2
    xp = np.pad(x, ((0, 0), (0, 0), (P, P), (P, P)))
    out = np.zeros((B, Co, H * W))
    for ky in range(K):
        for kx in range(K):
            patch = xp[:, :, ky:ky + H, kx:kx + W].reshape(B, Ci, H * W)
            # [Co,Ci] @ [B,Ci,HW]; contiguous operands so that numpy uses BLAS
            out += np.matmul(np.ascontiguousarray(w[:, :, ky, kx]), patch)
    out = out.reshape(B, Co, H, W) + b[None, :, None, None]
    return np.maximum(out, 0.0) if relu else out


def conv_bwd(x, w, out, dout, relu):
    """Returns (dx, dw, db) for the gradient dout at the (activated) output."""
    B, Ci, H, W = x.shape
    Co, _, K, _ = w.shape
    P = K // 2
    g = dout * (out > 0) if relu else dout
    xp = np.pad(x, ((0, 0), (0, 0), (P, P), (P, P)))
    dw = np.zeros_like(w, dtype=np.float64)
    dxp = np.zeros_like(xp, dtype=np.float64)
    g2 = np.ascontiguousarray(g, np.float64).reshape(B, Co, H * W)
    for ky in range(K):
        for kx in range(K):
            patch = xp[:, :, ky:ky + H, kx:kx + W].reshape(B, Ci, H * W)
            dw[:, :, ky, kx] = np.matmul(g2, patch.transpose(0, 2, 1)).sum(axis=0)
            wt = np.ascontiguousarray(w[:, :, ky, kx].T)
            dxp[:, :, ky:ky + H, kx:kx + W] += np.matmul(wt, g2).reshape(
                B, Ci, H, W)
    db = g.sum(axis=(0, 2, 3))
    dx = dxp[:, :, P:P + H, P:P + W]
    return dx, dw, db


def pool_fwd(x):
    B, C, H, W = x.shape
    return x.reshape(B, C, H // 2, 2, W // 2, 2).mean(axis=(3, 5))


def pool_bwd(dout):
    return np.repeat(np.repeat(dout, 2, axis=2), 2, axis=3) * 0.25


def up_fwd(x):
    return np.repeat(np.repeat(x, 2, axis=2), 2, axis=3)


def up_bwd(dout):
    B, C, H2, W2 = dout.shape
    return dout.reshape(B, C, H2 // 2, 2, W2 // 2, 2).sum(axis=(3, 5))


def mse(y, t):
    d = y - t
    return float(np.mean(d * d)), 2.0 * d / d.size


def unet_shapes(levels, base, cin=1, cout=1):
    """The parameter list of the U-Net, in storage order: (name, Co, Ci, K)."""
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


def unet_forward(params, x, levels):
    """params: dict name -> (w, b).  Returns (y, cache)."""
    cache = {}
    h = x
    skips = []
    for l in range(levels):
        cache[f"e{l}a_in"] = h
        a = conv_fwd(h, *params[f"e{l}a"], relu=True)
        cache[f"e{l}a_out"] = a
        e = conv_fwd(a, *params[f"e{l}b"], relu=True)
        cache[f"e{l}b_out"] = e
        skips.append(e)
        if l < levels - 1:
            h = pool_fwd(e)
    d = skips[-1]
    for l in range(levels - 2, -1, -1):
        cat = np.concatenate([up_fwd(d), skips[l]], axis=1)
        cache[f"d{l}a_in"] = cat
        a = conv_fwd(cat, *params[f"d{l}a"], relu=True)
        cache[f"d{l}a_out"] = a
        d = conv_fwd(a, *params[f"d{l}b"], relu=True)
        cache[f"d{l}b_out"] = d
    cache["out_in"] = d
    y = conv_fwd(d, *params["out"], relu=False)
    return y, cache


def unet_backward(params, cache, dy, levels):
    """Gradients dict name -> (dw, db) for the output gradient dy."""
    grads = {}
    gd, gw, gb = conv_bwd(cache["out_in"], params["out"][0], None, dy, relu=False)
    grads["out"] = (gw, gb)
    skip_grads = {}
    for l in range(0, levels - 1):
        ga, gw, gb = conv_bwd(cache[f"d{l}a_out"], params[f"d{l}b"][0], cache[f"d{l}b_out"], gd, True)
        grads[f"d{l}b"] = (gw, gb)
        gcat, gw, gb = conv_bwd(cache[f"d{l}a_in"], params[f"d{l}a"][0], cache[f"d{l}a_out"], ga, True)
        grads[f"d{l}a"] = (gw, gb)
        cu = params[f"d{l}a"][0].shape[1] - params[f"d{l}b"][0].shape[0]  # up-path channels
        skip_grads[l] = gcat[:, cu:]
        gd = up_bwd(gcat[:, :cu])
    ge = gd  # gradient at the bottom level's output
    for l in range(levels - 1, -1, -1):
        if l < levels - 1:
            ge = skip_grads[l] + ge
        ga, gw, gb = conv_bwd(cache[f"e{l}a_out"], params[f"e{l}b"][0], cache[f"e{l}b_out"], ge, True)
        grads[f"e{l}b"] = (gw, gb)
        gx, gw, gb = conv_bwd(cache[f"e{l}a_in"], params[f"e{l}a"][0], cache[f"e{l}a_out"], ga, True)
        grads[f"e{l}a"] = (gw, gb)
        if l > 0:
            ge = pool_bwd(gx)
    return grads
