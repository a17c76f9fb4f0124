"""Thin Python handle over a ``gvm_ctx`` (libgvm_hip C ABI) for harnesses.

All device memory is allocated through the library (``gvm_malloc``); all
computation runs in the HIP kernels of libgvm_hip.  Nothing here computes.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import native
from .problem import Problem


class DeviceBuffer:
    """A device allocation of ``nbytes`` made by gvm_malloc."""

    def __init__(self, nbytes: int):
        self.lib = native.load_hip()
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        native.check(self.lib.gvm_malloc(C.byref(p), self.nbytes), "gvm_malloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        native.check(self.lib.gvm_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes), "h2d")

    def download(self, dtype=np.float32, shape=None) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        native.check(self.lib.gvm_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes), "d2h")
        return out.reshape(shape) if shape is not None else out

    def free(self) -> None:
        if self.ptr:
            self.lib.gvm_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Graph:
    """A HIP graph captured on a context (gvm_capture_begin / gvm_capture_end)."""

    def __init__(self, ctx: "Context", handle: int):
        self.ctx = ctx
        self.h = handle

    def launch(self) -> None:
        native.check(self.ctx.lib.gvm_graph_launch(self.ctx.ctx, self.h), "gvm_graph_launch")

    def status_of_launch(self) -> int:
        return self.ctx.lib.gvm_graph_launch(self.ctx.ctx, self.h)

    def destroy(self) -> None:
        if self.h:
            self.ctx.lib.gvm_graph_destroy(self.h)
            self.h = None


class Context:
    """A gvm_ctx with one slice per call to :meth:`add`."""

    def __init__(self, problem: Problem, adjoint: str = "bf16x3", add_data: bool = True):
        self.lib = native.load_hip()
        self.problem = problem
        g = problem.geometry()
        ctx = C.c_void_p()
        noise = np.ascontiguousarray(problem.noise, np.float32)
        native.check(self.lib.gvm_create(C.byref(g), noise.ctypes.data, C.byref(ctx)), "gvm_create")
        self.ctx = ctx.value
        self.set_adjoint(adjoint)
        self.n = problem.M * problem.N
        if add_data:
            for f, uvw, vo, w in problem.all_slices():
                self.add(f, uvw, vo, w)

    def add(self, freq: float, uvw: np.ndarray, vo: np.ndarray, weight: np.ndarray) -> int:
        uvw = np.ascontiguousarray(uvw, np.float64)
        vo = np.ascontiguousarray(vo, np.float32)
        weight = np.ascontiguousarray(weight, np.float32)
        s = C.c_int()
        native.check(self.lib.gvm_add_visibilities(self.ctx, freq, uvw.shape[0], uvw.ctypes.data,
                                                   vo.ctypes.data, weight.ctypes.data, C.byref(s)),
                     "gvm_add_visibilities")
        return s.value

    def add_dev(self, freq: float, uvw_m: "DeviceBuffer", vo: "DeviceBuffer", weight: "DeviceBuffer",
                nvis: int) -> int:
        """A slice from device arrays in the reference's raw form (uvw in
        metres before the Hermitian flip), set up on the device."""
        s = C.c_int()
        native.check(self.lib.gvm_add_visibilities_dev(self.ctx, freq, nvis, uvw_m.ptr, vo.ptr,
                                                       weight.ptr, C.byref(s)),
                     "gvm_add_visibilities_dev")
        return s.value

    def set_adjoint(self, mode: str) -> None:
        native.check(self.lib.gvm_set_adjoint(self.ctx, native.ADJOINT_MODES[mode]), "gvm_set_adjoint")

    def image(self, img: np.ndarray) -> DeviceBuffer:
        assert img.dtype == np.float32 and img.size == 2 * self.n
        return DeviceBuffer.from_array(img)

    def zeros(self, count: int) -> DeviceBuffer:
        b = DeviceBuffer(4 * count)
        native.check(self.lib.gvm_vec_zero(self.ctx, b.ptr, count), "gvm_vec_zero")
        return b

    def chi2(self, img: DeviceBuffer, flag_opt: int = 0, normalize: bool = False) -> float:
        out = C.c_double()
        native.check(self.lib.gvm_chi2(self.ctx, img.ptr, flag_opt, int(normalize), C.byref(out)),
                     "gvm_chi2")
        return out.value

    def dchi2(self, img: DeviceBuffer, result: DeviceBuffer, flag_opt: int = 0,
              normalize: bool = False) -> None:
        native.check(self.lib.gvm_dchi2(self.ctx, img.ptr, result.ptr, flag_opt, int(normalize)),
                     "gvm_dchi2")

    def residuals(self, slice_index: int = 0, nvis: Optional[int] = None) -> np.ndarray:
        nvis = self.problem.nvis if nvis is None else nvis
        out = np.empty((nvis, 2), np.float32)
        native.check(self.lib.gvm_get_residuals(self.ctx, slice_index, out.ctypes.data),
                     "gvm_get_residuals")
        return out

    def prior_value(self, kind: str, img: DeviceBuffer, index: int = 0, prior: float = 0.001,
                    eta: float = -1.0) -> float:
        out = C.c_double()
        if kind == "entropy":
            st = self.lib.gvm_entropy_value(self.ctx, img.ptr, index, prior, eta, C.byref(out))
        elif kind == "tsv":
            st = self.lib.gvm_tsv_value(self.ctx, img.ptr, index, C.byref(out))
        else:
            st = self.lib.gvm_laplacian_value(self.ctx, img.ptr, index, C.byref(out))
        native.check(st, f"{kind} value")
        return out.value

    def prior_grad(self, kind: str, img: DeviceBuffer, dgi: DeviceBuffer, index: int = 0,
                   lam: float = 1.0, prior: float = 0.001, eta: float = -1.0) -> None:
        if kind == "entropy":
            st = self.lib.gvm_entropy_grad(self.ctx, img.ptr, dgi.ptr, index, prior, eta, lam)
        elif kind == "tsv":
            st = self.lib.gvm_tsv_grad(self.ctx, img.ptr, dgi.ptr, index, lam)
        else:
            st = self.lib.gvm_laplacian_grad(self.ctx, img.ptr, dgi.ptr, index, lam)
        native.check(st, f"{kind} gradient")

    def sync(self) -> None:
        native.check(self.lib.gvm_sync(self.ctx), "gvm_sync")

    # ---- HIP graphs of enqueued work ----
    def capture_begin(self) -> None:
        native.check(self.lib.gvm_capture_begin(self.ctx), "gvm_capture_begin")

    def capture_end(self) -> "Graph":
        g = C.c_void_p()
        native.check(self.lib.gvm_capture_end(self.ctx, C.byref(g)), "gvm_capture_end")
        return Graph(self, g.value)

    def chi2_enqueue(self, img: DeviceBuffer, slot: int, flag_opt: int = 0,
                     normalize: bool = False) -> None:
        native.check(self.lib.gvm_chi2_enqueue(self.ctx, img.ptr, flag_opt, int(normalize), slot),
                     "gvm_chi2_enqueue")

    def values_fetch_enqueue(self, n: int) -> None:
        native.check(self.lib.gvm_values_fetch_enqueue(self.ctx, n), "gvm_values_fetch_enqueue")

    def values_host(self, n: int) -> list[float]:
        out = (C.c_double * n)()
        native.check(self.lib.gvm_values_host(self.ctx, n, out), "gvm_values_host")
        return list(out)

    def set_fft_mode(self, mode: str) -> None:
        native.check(self.lib.gvm_set_fft_mode(self.ctx, {"auto": 0, "wide": 1, "narrow": 2}[mode]),
                     "gvm_set_fft_mode")

    def set_split_vis(self, n: int) -> None:
        native.check(self.lib.gvm_set_split_vis(self.ctx, n), "gvm_set_split_vis")

    def split_vis(self) -> int:
        v = C.c_long()
        native.check(self.lib.gvm_get_split_vis(self.ctx, C.byref(v)), "gvm_get_split_vis")
        return v.value

    def adjoint_computed_pixels(self) -> int:
        """gvm_adjoint_computed_pixels: the pixels whose sums the adjoint of
        the current form computes (the work of one launch is 4 x this x the
        visibilities)."""
        v = C.c_long()
        native.check(self.lib.gvm_adjoint_computed_pixels(self.ctx, C.byref(v)),
                     "gvm_adjoint_computed_pixels")
        return v.value

    def autotune(self, candidates=None) -> dict:
        """gvm_autotune_adjoint: picks the adjoint's range length by timing
        (a chi2 must have run).  Returns the choice, the candidates' times and
        whether the shape was already tuned."""
        if candidates is None:
            cand, nc = None, 6
        else:
            cand, nc = (C.c_long * len(candidates))(*candidates), len(candidates)
        ms = (C.c_double * nc)()
        chosen, cached = C.c_long(), C.c_int()
        native.check(self.lib.gvm_autotune_adjoint(self.ctx, cand, nc if cand else 0,
                                                   C.byref(chosen), ms, C.byref(cached)),
                     "gvm_autotune_adjoint")
        names = list(candidates) if candidates else [1536, 2048, 3072, 4096, 6144, 8192]
        # the default list stops at the longest range the image size allows
        # (3072 from 2048^2 on); the candidates not timed keep 0
        timed = {n: t for n, t in zip(names, list(ms)) if t > 0.0}
        return {"split_vis": chosen.value, "cached": bool(cached.value),
                "ms": None if cached.value else timed}

    def adjoint_time_ms(self) -> tuple[float, int]:
        ms = C.c_double()
        cnt = C.c_long()
        native.check(self.lib.gvm_timing_adjoint_ms(self.ctx, C.byref(ms), C.byref(cnt)), "timing")
        return ms.value, cnt.value

    def reset_timing(self) -> None:
        native.check(self.lib.gvm_timing_reset(self.ctx), "timing reset")

    def close(self) -> None:
        if self.ctx:
            self.lib.gvm_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# framework (libgvm): objective composed by factory id + optimisers
# ---------------------------------------------------------------------------
def fw_config(terms, lambdas=(), initial_values=(0.001, 0.0), eta=-1.0, flag_opt=0,
              nopositivity=False) -> native.FwConfig:
    """terms: list of (name, penalizator_index, prior) in the order of
    src/main.cu:199-203 (Chi2 first)."""
    cfg = native.FwConfig()
    cfg.nterms = len(terms)
    keep = []
    for k, (name, pen, prior) in enumerate(terms):
        b = name.encode()
        keep.append(b)
        cfg.terms[k] = native.FwTerm(b, pen, 0, 0, prior, 0)
    cfg.nlambdas = len(lambdas)
    for k, v in enumerate(lambdas):
        cfg.lambdas[k] = v
    cfg.initial_values[0], cfg.initial_values[1] = initial_values
    cfg.eta = eta
    cfg.nopositivity = int(nopositivity)
    cfg.flag_opt = flag_opt
    cfg._keep = keep  # keep the name buffers alive
    return cfg


class Objective:
    """An ObjectiveFunction built through the C++ factories (libgvm)."""

    def __init__(self, ctx: Context, cfg: native.FwConfig):
        self.fw = native.load_fw()
        self.ctx = ctx
        self.cfg = cfg
        h = C.c_void_p()
        native.check_fw(self.fw.gvm_fw_objective_create(ctx.ctx, C.byref(cfg), C.byref(h)),
                        "gvm_fw_objective_create")
        self.h = h.value

    def value(self, img: DeviceBuffer) -> float:
        v = C.c_double()
        native.check_fw(self.fw.gvm_fw_value(self.h, img.ptr, C.byref(v)), "gvm_fw_value")
        return v.value

    def gradient(self, img: DeviceBuffer, xi: DeviceBuffer, it: int) -> None:
        native.check_fw(self.fw.gvm_fw_gradient(self.h, img.ptr, xi.ptr, it), "gvm_fw_gradient")

    def term_values(self) -> list[float]:
        vals = (C.c_double * 8)()
        n = C.c_int()
        native.check_fw(self.fw.gvm_fw_term_values(self.h, vals, 8, C.byref(n)), "term values")
        return [vals[k] for k in range(n.value)]

    def optimize(self, img: DeviceBuffer, optimizer: str = "CG-FRPRMN", iterations: int = 50,
                 ftol: float = 1e-12, gtol: float = 1e-12, lbfgs_k: int = 0):
        res = native.FwResult()
        traj = (C.c_double * max(1, iterations))()
        native.check_fw(self.fw.gvm_fw_optimize(self.h, optimizer.encode(), iterations, ftol, gtol,
                                                lbfgs_k, img.ptr, C.byref(res), traj, iterations),
                        "gvm_fw_optimize")
        n = min(res.iterations, iterations)
        return res, [traj[k] for k in range(n)]

    def optimize_timed(self, img: DeviceBuffer, optimizer: str = "CG-FRPRMN", iterations: int = 50,
                       ftol: float = 1e-12, gtol: float = 1e-12, lbfgs_k: int = 0):
        """optimize() that also returns the wall-clock seconds from the start
        of the run to each trajectory entry."""
        res = native.FwResult()
        m = max(1, iterations)
        traj, secs = (C.c_double * m)(), (C.c_double * m)()
        native.check_fw(self.fw.gvm_fw_optimize_timed(self.h, optimizer.encode(), iterations, ftol,
                                                      gtol, lbfgs_k, img.ptr, C.byref(res), traj,
                                                      secs, iterations), "gvm_fw_optimize_timed")
        n = min(res.iterations, iterations)
        return res, [traj[k] for k in range(n)], [secs[k] for k in range(n)]

    def set_graphs(self, on: bool) -> None:
        """Replay captured HIP graphs of repeated evaluations."""
        native.check_fw(self.fw.gvm_fw_set_graphs(self.h, int(on)), "gvm_fw_set_graphs")

    def set_fused_terms(self, on: bool) -> None:
        """Priors after Chi2 in one pass (default) or term by term."""
        native.check_fw(self.fw.gvm_fw_set_fused_terms(self.h, int(on)), "gvm_fw_set_fused_terms")

    def set_lbfgs_host_scalars(self, on: bool) -> None:
        """CG-LBFGS reads every dot product of its recursion back to the host
        (the reference's way) instead of keeping the scalars on the device."""
        native.check_fw(self.fw.gvm_fw_set_lbfgs_host_scalars(self.h, int(on)),
                        "gvm_fw_set_lbfgs_host_scalars")

    def set_line_search_replay(self, xmins) -> None:
        """The next optimize() takes its line-search step lengths from xmins
        in turn instead of searching (an empty list restores the search)."""
        arr = np.ascontiguousarray(np.asarray(xmins, np.float64))
        ptr = arr.ctypes.data_as(C.POINTER(C.c_double)) if arr.size else None
        native.check_fw(self.fw.gvm_fw_set_line_search_replay(self.h, ptr, int(arr.size)),
                        "gvm_fw_set_line_search_replay")

    def line_search_steps(self) -> list[float]:
        """The step lengths the last optimize() applied, in order."""
        n = C.c_int()
        native.check_fw(self.fw.gvm_fw_line_search_steps(self.h, None, 0, C.byref(n)),
                        "gvm_fw_line_search_steps")
        out = (C.c_double * max(1, n.value))()
        native.check_fw(self.fw.gvm_fw_line_search_steps(self.h, out, n.value, C.byref(n)),
                        "gvm_fw_line_search_steps")
        return [out[k] for k in range(n.value)]

    def graph_launches(self) -> int:
        n = C.c_long()
        native.check_fw(self.fw.gvm_fw_graph_launches(self.h, C.byref(n)), "gvm_fw_graph_launches")
        return n.value

    def set_flag(self, flag_opt: int) -> None:
        """Optimizer::setFlag for the following calls (schedule flag)."""
        native.check_fw(self.fw.gvm_fw_set_flag(self.h, flag_opt), "gvm_fw_set_flag")

    def optimize_schedule(self, img: DeviceBuffer, flags=(0, 1, 2, 3), **kw):
        """MFS::run's schedule (src/mfs.cu:995-1009): the optimiser runs once
        per flag, in order, on the same image.  Returns [(res, traj)]."""
        out = []
        for f in flags:
            self.set_flag(f)
            out.append(self.optimize(img, **kw))
        return out

    def close(self) -> None:
        if self.h:
            self.fw.gvm_fw_objective_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
