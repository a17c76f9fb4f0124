"""ctypes wrapper of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, always to check results against (or as the
timed CPU baseline), never by the product path.  Every function restates the reference
(file:line citations are in gvm_oracle.c).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


class OCplx(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class OGeom(C.Structure):
    _fields_ = [
        ("M", C.c_int), ("N", C.c_int),
        ("deltax", C.c_double), ("deltay", C.c_double),
        ("deltau", C.c_double), ("deltav", C.c_double),
        ("x0_ref", C.c_float), ("y0_ref", C.c_float),
        ("x0_phs", C.c_float), ("y0_phs", C.c_float),
        ("fg_scale", C.c_float), ("noise_cut", C.c_float), ("minpix", C.c_float),
        ("eta", C.c_float), ("threshold", C.c_float), ("nu_0", C.c_float),
        ("noise", C.c_void_p), ("gcf", C.c_void_p),
    ]


class OTaper(C.Structure):
    _fields_ = [("amplitude", C.c_float), ("u_0", C.c_double), ("v_0", C.c_double),
                ("sigma_maj", C.c_float), ("sigma_min", C.c_float), ("bpa", C.c_float)]


class OBeam(C.Structure):
    _fields_ = [("kind", C.c_int), ("diameter", C.c_double), ("pb_factor", C.c_double),
                ("pb_cutoff", C.c_double)]


class OChan(C.Structure):
    _fields_ = [("freq", C.c_float), ("nvis", C.c_long), ("uvw", C.c_void_p),
                ("vo", C.c_void_p), ("weight", C.c_void_p)]


class OFi(C.Structure):
    _fields_ = [("kind", C.c_int), ("lam", C.c_double), ("imageIndex", C.c_int),
                ("imageToAdd", C.c_int), ("prior", C.c_double), ("eta", C.c_double),
                ("iteration", C.c_int), ("value", C.c_double)]


class OObjective(C.Structure):
    _fields_ = [
        ("g", C.POINTER(OGeom)), ("b", C.POINTER(OBeam)), ("ch", C.POINTER(OChan)),
        ("nch", C.c_int), ("image_count", C.c_int), ("flag_opt", C.c_int),
        ("normalize", C.c_int), ("separable", C.c_int), ("nopositivity", C.c_int),
        ("nfi", C.c_int), ("fi", OFi * 8), ("vr", C.c_void_p), ("dphi", C.c_void_p),
        ("dgi", C.c_void_p), ("n_fcalls", C.c_long), ("n_gcalls", C.c_long),
    ]


class OResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("n_values", C.c_long), ("n_gradients", C.c_long),
                ("exit_reason", C.c_int), ("fret", C.c_double),
                ("final_value", C.c_double), ("n_fcalls", C.c_long), ("n_gcalls", C.c_long)]


F1 = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
VALUE_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.POINTER(C.c_double))
GRAD_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int)


class OProblem(C.Structure):
    _fields_ = [("n", C.c_long), ("image_count", C.c_int), ("nopositivity", C.c_int),
                ("floor_v", C.c_double), ("value", C.c_void_p), ("gradient", C.c_void_p),
                ("ctx", C.c_void_p)]
FI_KIND = {"Chi2": 0, "Entropy": 1, "TotalSquaredVariation": 2, "Laplacian": 3}

_lib = None


def build() -> None:
    """Compiles liboracle.so with oracle/Makefile."""
    subprocess.run(["make", "-C", str(HERE), "-s"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P, D, I = C.c_void_p, C.c_double, C.c_int
        GP, BP, CP = C.POINTER(OGeom), C.POINTER(OBeam), C.POINTER(OChan)
        sig = {
            "orc_fft2d": ([P, I, I, I], I),
            "orc_clip2IWNoise": ([GP, P, I], None),
            "orc_model_grid": ([GP, BP, P, C.c_float, P], None),
            "orc_degrid_chi2": ([GP, P, CP, P, P], D),
            "orc_chi2": ([GP, BP, CP, I, P, I, I, P], D),
            "orc_dchi2_channel": ([GP, BP, CP, P, I, I, P], None),
            "orc_dchi2_direct_pixels": ([GP, BP, CP, P, I, P, I, P], None),
            "orc_dchi2": ([GP, BP, CP, I, P, P, P, I, I, I], None),
            "orc_dchi2_pixels": ([GP, BP, CP, I, P, P, I, I, P, I, P], None),
            "orc_set_threads": ([I], None),
            "orc_max_threads": ([], I),
            "orc_entropy_value": ([GP, P, I, D, D], D),
            "orc_entropy_grad": ([GP, P, I, D, D, D, P], None),
            "orc_tsv_value": ([GP, P, I], D),
            "orc_tsv_grad": ([GP, P, I, D, P], None),
            "orc_laplacian_value": ([GP, P, I], D),
            "orc_laplacian_grad": ([GP, P, I, D, P], None),
            "orc_attenuation": ([BP, D, I, I, C.c_float, C.c_float, D, D], D),
            "orc_pb_image": ([BP, D, I, I, C.c_float, C.c_float, D, D, P], None),
            "orc_hermitian_to_lambda": ([C.c_long, P, P, C.c_float], None),
            "orc_beam_and_noise": ([CP, I, D, P, P, P, P, P], None),
            "orc_noise_jypix": ([D, D, D, D], D),
            "orc_noise_image": ([BP, C.c_float, I, I, D, D, C.c_float, C.c_float, D, P, P], None),
            "orc_of_init": ([C.POINTER(OObjective)], I),
            "orc_of_free": ([C.POINTER(OObjective)], None),
            "orc_of_calc_function": ([C.POINTER(OObjective), P], D),
            "orc_of_calc_gradient": ([C.POINTER(OObjective), P, P, I], None),
            "orc_mnbrak": ([P, P, P, P, P, P, F1, P], None),
            "orc_brent": ([D, D, D, D, P, F1, P], D),
            "orc_of_problem": ([C.POINTER(OObjective)], OProblem),
            "orc_linmin": ([C.POINTER(OProblem), P, P], D),
            "orc_set_xmin_log": ([P, I], None),
            "orc_xmin_log_count": ([], I),
            "orc_frprmn": ([C.POINTER(OProblem), P, I, D, D, C.POINTER(OResult), P, I], I),
            "orc_lbfgs": ([C.POINTER(OProblem), P, I, I, D, D, C.POINTER(OResult), P, I], I),
            "orc_do_gridding_pillbox": ([I, I, D, D, CP, I, I, I, P, P, P], None),
            "orc_grid_visibilities": ([I, I, D, D, C.c_float, C.c_long, P, P, P, I, I, I, I, P,
                                       P, P, P, P], None),
            "orc_weighting": ([I, C.c_float, C.POINTER(OTaper), I, I, D, D, I, P, P, P, P, I],
                              None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def grid_visibilities(M: int, N: int, deltax: float, deltay: float, freq: float, uvw_m, vo, weight,
                      taps, support_x: int, support_y: int):
    """do_gridding of one slice (gvm_oracle_grid.c): returns the gridded
    (uvw metres (n, 3) float64, vo (n, 2) float32, weight (n,) float32).
    taps: (m, n) float32 kernel."""
    uvw_m = np.ascontiguousarray(uvw_m, np.float64)
    vo = np.ascontiguousarray(vo, np.float32)
    weight = np.ascontiguousarray(weight, np.float32)
    taps = np.ascontiguousarray(taps, np.float32)
    cap = M * N
    ou = np.zeros((cap, 3), np.float64)
    ov = np.zeros((cap, 2), np.float32)
    ow = np.zeros(cap, np.float32)
    nout = C.c_long()
    lib().orc_grid_visibilities(M, N, deltax, deltay, freq, uvw_m.shape[0], _ptr(uvw_m), _ptr(vo),
                                _ptr(weight), taps.shape[0], taps.shape[1], support_x, support_y,
                                _ptr(taps), _ptr(ou), _ptr(ov), _ptr(ow), C.byref(nout))
    n = nout.value
    return ou[:n].copy(), ov[:n].copy(), ow[:n].copy()


WEIGHTING = {"natural": 0, "uniform": 1, "briggs": 2, "radial": 3}


def weighting(scheme: str, slices, M: int, N: int, deltax: float, deltay: float,
              robust: float = 0.0, taper=None, float_totals: bool = True):
    """WeightingScheme::apply over slices [(freq, uvw_m, weight)]: returns
    the new weight arrays (float32).  taper: dict of OTaper fields or None."""
    nsl = len(slices)
    freqs = (C.c_float * nsl)(*[s[0] for s in slices])
    nv = (C.c_long * nsl)(*[s[1].shape[0] for s in slices])
    us = [np.ascontiguousarray(s[1], np.float64) for s in slices]
    ws = [np.ascontiguousarray(s[2], np.float32).copy() for s in slices]
    up = (C.c_void_p * nsl)(*[_ptr(u) for u in us])
    wp = (C.c_void_p * nsl)(*[_ptr(w) for w in ws])
    t = OTaper(**taper) if taper is not None else None
    lib().orc_weighting(WEIGHTING[scheme], robust, C.byref(t) if t is not None else None, M, N,
                        deltax, deltay, nsl, freqs, nv, up, wp, int(float_totals))
    return ws


def fft2d(a: np.ndarray, sign: int = +1) -> np.ndarray:
    """Unnormalised 2-D DFT with e^{sign 2 pi i} (power-of-two sizes)."""
    x = np.ascontiguousarray(a, np.complex128).copy()
    rc = lib().orc_fft2d(_ptr(x), x.shape[0], x.shape[1], sign)
    if rc != 0:
        raise ValueError("oracle FFT needs power-of-two sizes")
    return x


def mnbrak(f, ax: float, bx: float):
    cb = F1(lambda x, _: float(f(x)))
    v = [C.c_double(ax), C.c_double(bx), C.c_double(), C.c_double(), C.c_double(), C.c_double()]
    lib().orc_mnbrak(*[C.addressof(t) for t in v], cb, None)
    return tuple(t.value for t in v)


def brent(f, ax: float, bx: float, cx: float, tol: float = 1e-7):
    cb = F1(lambda x, _: float(f(x)))
    xmin = C.c_double()
    fx = lib().orc_brent(ax, bx, cx, tol, C.addressof(xmin), cb, None)
    return xmin.value, fx


class Oracle:
    """The hot path of one problem (single channel) in FP64.

    ``noise`` defaults to the problem's float32 noise image so the mask is
    bit-identical to the one the device uses."""

    def __init__(self, problem, separable: bool = True, noise=None, slices=None, gcf=None):
        """slices: optional list of (freq, uvw, vo, weight) replacing the
        problem's single channel; gcf: optional gridding correction image."""
        self.p = problem
        L = lib()
        self.noise = np.ascontiguousarray(problem.noise if noise is None else noise, np.float32)
        self.gcf = None if gcf is None else np.ascontiguousarray(gcf, np.float32)
        dxr = problem.deltax * np.pi / 180.0
        dyr = problem.deltay * np.pi / 180.0
        self.geom = OGeom(problem.M, problem.N, problem.deltax, problem.deltay,
                          1.0 / (problem.M * dxr), 1.0 / (problem.N * dyr),
                          problem.x0, problem.y0, problem.x0, problem.y0,
                          problem.fg_scale, problem.noise_cut, problem.minpix, problem.eta,
                          problem.threshold, problem.nu_0, _ptr(self.noise),
                          _ptr(self.gcf) if self.gcf is not None else None)
        self.beam = OBeam(problem.beam_kind, problem.antenna_diameter, problem.pb_factor,
                          problem.pb_cutoff)
        if slices is None:
            slices = (problem.all_slices() if hasattr(problem, "all_slices")
                      else [(problem.freq, problem.uvw, problem.vo, problem.weight)])
        self._keep = []
        chans = []
        for f, u, v, w in slices:
            u = np.ascontiguousarray(u, np.float64)
            v = np.ascontiguousarray(v, np.float32)
            w = np.ascontiguousarray(w, np.float32).copy()
            self._keep.append((u, v, w))
            chans.append(OChan(f, u.shape[0], _ptr(u), _ptr(v), _ptr(w)))
        self.nch = len(chans)
        self.chan = (OChan * self.nch)(*chans)
        self.uvw, self.vo, self.weight = self._keep[0]
        self.ntot = sum(k[0].shape[0] for k in self._keep)
        self.separable = separable
        self.n = problem.M * problem.N
        self._lib = L

    # ---- forward ----
    def chi2(self, image, flag_opt: int = 0, normalize: bool = False):
        """Returns (chi2, residuals complex128, clipped image float64)."""
        img = np.ascontiguousarray(image, np.float64).reshape(2, self.p.N, self.p.M).copy()
        vr = np.zeros(self.ntot, np.complex128)
        v = self._lib.orc_chi2(C.byref(self.geom), C.byref(self.beam), self.chan, self.nch,
                               _ptr(img), flag_opt, int(normalize), _ptr(vr))
        return v, vr, img

    def model_grid(self, image):
        img = np.ascontiguousarray(image, np.float64).reshape(2, self.p.N, self.p.M).copy()
        grid = np.zeros((self.p.N, self.p.M), np.complex128)
        self._lib.orc_model_grid(C.byref(self.geom), C.byref(self.beam), _ptr(img), self.p.freq,
                                 _ptr(grid))
        return grid

    def pb(self, freq=None):
        out = np.zeros((self.p.N, self.p.M), np.float64)
        self._lib.orc_pb_image(C.byref(self.beam), self.p.freq if freq is None else freq,
                               self.p.M, self.p.N, self.p.x0, self.p.y0, self.p.deltax,
                               self.p.deltay, _ptr(out))
        return out

    # ---- adjoint ----
    def dchi2(self, image, vr, flag_opt: int = 0, normalize: bool = False, separable=None):
        img = np.ascontiguousarray(image, np.float64).reshape(2, self.p.N, self.p.M)
        vr = np.ascontiguousarray(vr, np.complex128)
        res = np.zeros((2, self.p.N, self.p.M), np.float64)
        sep = self.separable if separable is None else separable
        self._lib.orc_dchi2(C.byref(self.geom), C.byref(self.beam), self.chan, self.nch,
                            _ptr(img), _ptr(vr), _ptr(res), flag_opt, int(normalize), int(sep))
        return res

    def dchi2_channel(self, vr, normalize: bool = False, separable=None):
        vr = np.ascontiguousarray(vr, np.complex128)
        out = np.zeros((self.p.N, self.p.M), np.float64)
        sep = self.separable if separable is None else separable
        self._lib.orc_dchi2_channel(C.byref(self.geom), C.byref(self.beam), self.chan, _ptr(vr),
                                    int(normalize), int(sep), _ptr(out))
        return out

    def dchi2_pixels(self, vr, pixels, normalize: bool = False):
        vr = np.ascontiguousarray(vr, np.complex128)
        pix = np.ascontiguousarray(np.asarray(pixels, np.int32).reshape(-1, 2))
        out = np.zeros(pix.shape[0], np.float64)
        self._lib.orc_dchi2_direct_pixels(C.byref(self.geom), C.byref(self.beam), self.chan,
                                          _ptr(vr), int(normalize), _ptr(pix), pix.shape[0],
                                          _ptr(out))
        return out

    def dchi2_pixels_all(self, image, vr, pixels, flag_opt: int = 0, normalize: bool = False):
        """dchi2 over every channel (direct sums, chain rule to I_nu_0 or
        alpha) at sampled pixels: array [npix, 2]."""
        img = np.ascontiguousarray(image, np.float64).reshape(2, self.p.N, self.p.M)
        vr = np.ascontiguousarray(vr, np.complex128)
        pix = np.ascontiguousarray(np.asarray(pixels, np.int32).reshape(-1, 2))
        out = np.zeros((pix.shape[0], 2), np.float64)
        self._lib.orc_dchi2_pixels(C.byref(self.geom), C.byref(self.beam), self.chan, self.nch,
                                   _ptr(img), _ptr(vr), flag_opt, int(normalize), _ptr(pix),
                                   pix.shape[0], _ptr(out))
        return out

    # ---- priors ----
    def prior_value(self, kind: str, image, index: int = 0, prior: float = 0.001,
                    eta: float = -1.0) -> float:
        img = np.ascontiguousarray(image, np.float64)
        g = C.byref(self.geom)
        if kind == "entropy":
            return self._lib.orc_entropy_value(g, _ptr(img), index, prior, eta)
        if kind == "tsv":
            return self._lib.orc_tsv_value(g, _ptr(img), index)
        return self._lib.orc_laplacian_value(g, _ptr(img), index)

    def prior_grad(self, kind: str, image, index: int = 0, lam: float = 1.0,
                   prior: float = 0.001, eta: float = -1.0):
        img = np.ascontiguousarray(image, np.float64)
        out = np.zeros((self.p.N, self.p.M), np.float64)
        g = C.byref(self.geom)
        if kind == "entropy":
            self._lib.orc_entropy_grad(g, _ptr(img), index, prior, eta, lam, _ptr(out))
        elif kind == "tsv":
            self._lib.orc_tsv_grad(g, _ptr(img), index, lam, _ptr(out))
        else:
            self._lib.orc_laplacian_grad(g, _ptr(img), index, lam, _ptr(out))
        return out

    # ---- objective + optimisers ----
    def objective(self, terms, lambdas=(), flag_opt: int = 0, nopositivity: bool = False):
        """terms: list of (name, penalizator_index, prior), like
        gpuvmem_amd.device.fw_config."""
        of = OObjective()
        of.g = C.pointer(self.geom)
        of.b = C.pointer(self.beam)
        of.ch = C.cast(self.chan, C.POINTER(OChan))
        of.nch = self.nch
        of.image_count = 2
        of.flag_opt = flag_opt
        of.normalize = 0
        of.separable = int(self.separable)
        of.nopositivity = int(nopositivity)
        k = 0
        for name, pen, prior in terms:
            if name == "Chi2":
                lam = 1.0 if pen == -1 else lambdas[pen]
            else:
                lam = lambdas[pen] if 0 <= pen < len(lambdas) else 0.0
            if lam == 0.0:
                continue  # ObjectiveFunction::addFi drops zero-lambda terms
            of.fi[k] = OFi(FI_KIND[name], lam, 0, 0, prior, -1.0, 0, 0.0)
            k += 1
        of.nfi = k
        rc = self._lib.orc_of_init(C.byref(of))
        if rc != 0:
            raise MemoryError("oracle objective allocation failed")
        return of

    def free(self, of) -> None:
        self._lib.orc_of_free(C.byref(of))

    def value(self, of, image) -> tuple[float, np.ndarray]:
        img = np.ascontiguousarray(image, np.float64).copy()
        v = self._lib.orc_of_calc_function(C.byref(of), _ptr(img))
        return v, img

    def gradient(self, of, image, it: int) -> np.ndarray:
        img = np.ascontiguousarray(image, np.float64).copy()
        xi = np.zeros_like(img)
        self._lib.orc_of_calc_gradient(C.byref(of), _ptr(img), _ptr(xi), it)
        return xi

    def optimize(self, of, image, optimizer: str = "CG-FRPRMN", iterations: int = 50,
                 ftol: float = 1e-12, gtol: float = 1e-12, lbfgs_k: int = 100,
                 steps: bool = False):
        """Returns (final image, OResult, values after each line search), and
        with steps=True also the step length linmin accepted at each line
        search."""
        img = np.ascontiguousarray(image, np.float64).copy()
        pr = self._lib.orc_of_problem(C.byref(of))
        if not steps:
            return run_optimizer(pr, img, optimizer, iterations, ftol, gtol, lbfgs_k)
        log = np.zeros(max(1, iterations + 1), np.float64)
        self._lib.orc_set_xmin_log(_ptr(log), len(log))
        try:
            out = run_optimizer(pr, img, optimizer, iterations, ftol, gtol, lbfgs_k)
            count = int(self._lib.orc_xmin_log_count())
        finally:
            self._lib.orc_set_xmin_log(None, 0)
        return (*out, log[: min(count, len(log))].copy())


    def optimize_schedule(self, of, image, flags=(0, 1, 2, 3), **kw):
        """MFS::run's schedule (src/mfs.cu:995-1009): one optimiser run per
        schedule flag, in order, on the same image.  Returns (final image,
        [(OResult, trajectory)])."""
        img = np.ascontiguousarray(image, np.float64).copy()
        out = []
        for f in flags:
            of.flag_opt = f
            img, res, traj = self.optimize(of, img, **kw)
            out.append((res, traj))
        return img, out


def run_optimizer(pr: OProblem, img: np.ndarray, optimizer: str, iterations: int,
                  ftol: float = 1e-12, gtol: float = 1e-12, lbfgs_k: int = 100):
    """Runs frprmn ("CG-FRPRMN") or lbfgs ("CG-LBFGS") on img in place."""
    res = OResult()
    traj = np.zeros(max(1, iterations), np.float64)
    L = lib()
    if optimizer == "CG-FRPRMN":
        L.orc_frprmn(C.byref(pr), _ptr(img), iterations, ftol, gtol, C.byref(res), _ptr(traj),
                     iterations)
    elif optimizer == "CG-LBFGS":
        L.orc_lbfgs(C.byref(pr), _ptr(img), lbfgs_k, iterations, ftol, gtol, C.byref(res),
                    _ptr(traj), iterations)
    else:
        raise KeyError(optimizer)
    return img, res, traj[: min(res.iterations, iterations)].copy()


class PyProblem:
    """An oracle optimiser problem defined by Python callbacks (tests):
    value(p) -> float and gradient(p, it) -> array, on image_count images of
    n values (numpy views of the oracle's buffers)."""

    def __init__(self, n, image_count, value, gradient, nopositivity=True, floor_v=0.0):
        self.n, self.ic = n, image_count
        self._value, self._gradient = value, gradient
        tot = n * image_count

        def v(_, p):
            return float(self._value(np.ctypeslib.as_array(p, (tot,))))

        def g(_, p, xi, it):
            out = np.ctypeslib.as_array(xi, (tot,))
            out[:] = self._gradient(np.ctypeslib.as_array(p, (tot,)), it)

        self._v = VALUE_FN(v)
        self._g = GRAD_FN(g)
        self.problem = OProblem(n, image_count, int(nopositivity), floor_v,
                                C.cast(self._v, C.c_void_p), C.cast(self._g, C.c_void_p), None)

    def optimize(self, x0, optimizer="CG-FRPRMN", iterations=50, ftol=1e-12, gtol=1e-12,
                 lbfgs_k=100):
        img = np.ascontiguousarray(x0, np.float64).copy()
        return run_optimizer(self.problem, img, optimizer, iterations, ftol, gtol, lbfgs_k)


def set_threads(n: int) -> None:
    """Threads of the oracle's OpenMP loops from now on (also for a runtime
    that was loaded before, when OMP_NUM_THREADS is no longer read)."""
    os.environ["OMP_NUM_THREADS"] = str(n)
    lib().orc_set_threads(n)
