"""Objective + gradient evaluations per second on one MI355X.

Workload (the headline configuration): 1024x1024 image, 0.05" cells, 50
antennas (Gaussian-radial in a 2 km disk), 4 h at 10 s dumps ->
1,764,000 visibilities at 230 GHz, one channel; sky of 200 Gaussian
components; objective chi2 + Entropy (lambda 0.01, G 0.001) + TSV
(lambda 1e-3), composed through the C++ factories.  One step = one
ObjectiveFunction::calcFunction + one calcGradient (the pair of
src/frprmn.cu:130-135) at a fixed image, all inputs resident in HBM.

Prints one JSON line.  The roofline figure is for the dominant kernel, the
DChi2 adjoint GEMM: algorithmic work 4 flop per (pixel, visibility) over the
pixels it computes (its 128 x 64 sub-tiles that hold a pixel on the mask; the
reference's DChi2 computes the pixels on the mask) per launch, over its
average duration (HIP events on the context's stream).  The cpu_baseline leg
times the FP64 oracle on the host cores on a bounded sample (full-size
forward chi2 + the separable FP64 gradient over a slice of the
visibilities, extrapolated linearly in the visibility count).  The
time_to_image object times L-BFGS to a common objective with the plain FP32
and the split-precision adjoint, and the split form once more from the
learned warm start of the committed network (gpuvmem_amd/data/).

Other modes, each printing its own JSON line: --workload 0/2/3 (other
workloads), --time-to-image ITERS (split precision against plain FP32 at
equal objective, median over observations), --precision-study, --gridding,
--warm-start (the learned warm start; --ws-* options), --graphs and
--autotune (modifiers of the evals/s run).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# the learned warm start of the 1024^2 workload (trained by bench.py --warm-start
# --ws-workload 1 --ws-save-model)
WARM_MODEL = ROOT / "gpuvmem_amd" / "data" / "warm_start_1024sq.npz"
PEAK_TFLOPS = {"f16x3": 2500.0, "bf16x3": 2500.0, "f32": 157.3, "direct": 157.3}
DTYPE = {
    "f16x3": "fp16 hi/lo x3 MFMA adjoint (fp32 accumulate, fp64 combine); fp32 forward",
    "bf16x3": "bf16 hi/lo x3 MFMA adjoint (fp32 accumulate, fp64 combine); fp32 forward",
    "f32": "fp32 MFMA adjoint; fp32 forward",
    "direct": "fp32 VALU direct-sum adjoint; fp32 forward",
}


def bench_image(p) -> np.ndarray:
    """A fixed, structured image: the true sky in internal units plus the
    positivity floor (alpha = 0)."""
    img = p.initial_image()
    img[0] = np.maximum(p.sky / p.fg_scale * 0.5, p.minpix).astype(np.float32)
    return img


def host_cpus() -> dict:
    """The host's processor count (nproc: the CPUs this process may run on),
    os.cpu_count(), the CPU quota of the process' cgroup (cpu.max, if any),
    the processor model (lscpu) and the thread count the CPU legs use: nproc,
    capped by the quota when there is one."""
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    quota = None
    try:
        q, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    model = None
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    threads = nproc if quota is None else max(1, min(nproc, int(quota)))
    return {"nproc": nproc, "os_cpu_count": os.cpu_count(), "cgroup_cpu_quota": quota,
            "cpu_model": model, "threads": threads}


def cpu_baseline(p, hc: dict, target_s: float) -> dict:
    """The FP64 oracle on the host cores: full-size forward chi2, the
    separable FP64 gradient on a slice of the visibilities (extrapolated
    linearly in the visibility count) and the reference's direct-sum
    gradient (src/functions.cu:3698-3791) on a sample of masked pixels
    (extrapolated linearly in the masked pixel count).  `value` uses the
    faster gradient form."""
    threads = hc["threads"]
    sys.path.insert(0, str(ROOT / "oracle"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import oracle

    oracle.set_threads(threads)

    img = bench_image(p).astype(np.float64)
    orc = oracle.Oracle(p, separable=True)
    t0 = time.perf_counter()
    _, vr, clipped = orc.chi2(img)
    t_fwd = time.perf_counter() - t0
    # gradient on a slice of the visibilities (of the first channel); grow the
    # slice until it costs about half of target_s (the growing rounds take the
    # rest), then extrapolate linearly in the total visibility count
    n = p.nvis_total
    frac_n = max(256, p.nvis // 512)
    while True:
        sub = oracle.Oracle(_slice_problem(p, frac_n), separable=True)
        t0 = time.perf_counter()
        sub.dchi2_channel(vr[:frac_n])
        dt = time.perf_counter() - t0
        if dt > target_s / 2 or frac_n >= p.nvis:
            break
        frac_n = min(p.nvis, int(frac_n * max(2.0, (target_s / 2) / max(dt, 1e-3))))
    t_grad = dt * n / frac_n
    # direct sums (one sincos per pixel and visibility) on masked pixels, a
    # multiple of the thread count, grown to about a quarter of target_s
    inside = np.argwhere(p.noise < p.noise_cut)
    rng = np.random.default_rng(3)
    npx = threads
    while True:
        pix = inside[rng.choice(len(inside), min(npx, len(inside)), replace=False)]
        t0 = time.perf_counter()
        orc.dchi2_pixels(vr, pix)
        dtd = time.perf_counter() - t0
        if dtd > target_s / 4 or npx >= len(inside):
            break
        npx = min(len(inside), threads * max(2 * (npx // threads),
                                             int(npx / threads * (target_s / 4) / max(dtd, 1e-3))))
    npx = len(pix)
    t_direct = dtd * len(inside) / npx
    evals = 1.0 / (t_fwd + min(t_grad, t_direct))
    grid_s, grid_k = gridding_comparator(p, oracle, threads)
    return {
        "value": evals,
        "unit": "evals/s",
        "cores": threads,
        "kind": "port",
        **{k: hc[k] for k in ("nproc", "os_cpu_count", "cgroup_cpu_quota", "cpu_model")},
        "sample": (f"FP64 oracle on {threads} threads: full-size forward chi2 ({t_fwd:.2f} s) + "
                   f"separable FP64 gradient over {frac_n} of {n} visibilities ({dt:.2f} s), "
                   f"extrapolated x{n / frac_n:.1f} to {t_grad:.1f} s; value uses the faster of "
                   f"that and the direct sum"),
        "gradient_seconds": {"separable_fp64": t_grad, "direct_sum_fp64": t_direct},
        "direct_sum": {
            "evals_per_s": 1.0 / (t_fwd + t_direct),
            "sample": (f"direct FP64 sums (the reference's DChi2 algorithm) over all {n} "
                       f"visibilities at {npx} of {len(inside)} masked pixels ({dtd:.2f} s), "
                       f"extrapolated x{len(inside) / npx:.0f} to {t_direct:.0f} s"),
        },
        "gridding_comparator": {
            "seconds_per_grid": grid_s,
            "threads": threads,
            "sample": (f"CPU restatement of the OpenMP gridding path do_gridding "
                       f"(src/functions.cu:1339-1653), 7x7 pillbox, {grid_k} of {n} visibilities "
                       f"(both Hermitian halves), extrapolated linearly"),
        },
    }


def gridding_comparator(p, oracle, threads: int, k: int = 100000):
    """Times the CPU gridding comparator on k visibilities, returns the time
    extrapolated to the whole observation and k."""
    import ctypes as C

    k = min(k, p.nvis)
    lam = float(np.float32(2.99792458e8) / np.float32(p.freq))
    uvw = np.ascontiguousarray(p.uvw[:k] * lam)  # metres, as the gridder reads them
    vo = np.ascontiguousarray(p.vo[:k])
    w = np.ascontiguousarray(p.weight[:k])
    ch = (oracle.OChan * 1)(oracle.OChan(p.freq, k, uvw.ctypes.data, vo.ctypes.data, w.ctypes.data))
    n = p.M * p.N
    re, im, gw = np.zeros(n), np.zeros(n), np.zeros(n)
    du = 1.0 / (p.M * p.deltax * np.pi / 180.0)
    dv = 1.0 / (p.N * p.deltay * np.pi / 180.0)
    t0 = time.perf_counter()
    oracle.lib().orc_do_gridding_pillbox(p.M, p.N, du, dv, ch, 1, 3, threads, re.ctypes.data,
                                         im.ctypes.data, gw.ctypes.data)
    return (time.perf_counter() - t0) * p.nvis_total / k, k


def bench_gridding(p, steps: int, warmup: int, threads: int, adjoint: str) -> dict:
    """The gridded-visibility branch of the set-up on the device (Briggs
    weighting, then gridding) timed on the workload's rows, beside the CPU
    restatement of the reference's OpenMP gridder, and the objective step on
    the gridded rows."""
    from gpuvmem_amd.device import Context, DeviceBuffer, Objective, fw_config
    from gpuvmem_amd.gridding import GriddingKernel, add_gridded_dev, apply_weighting_dev
    from gpuvmem_amd.gridding import grid_visibilities_dev

    lam = float(np.float32(2.99792458e8) / np.float32(p.freq))
    uvw = np.ascontiguousarray(p.uvw * lam)  # metres (the flip does not change the gridding)
    n = uvw.shape[0]
    ctx = Context(p, adjoint=adjoint, add_data=False)
    bu, bv = DeviceBuffer.from_array(uvw), DeviceBuffer.from_array(p.vo)
    bw = DeviceBuffer.from_array(p.weight)
    out = {}
    # Briggs weighting (robust 0): two passes over the rows with sorted cells
    ts = []
    for _ in range(3):
        bw.upload(p.weight)
        t0 = time.perf_counter()
        apply_weighting_dev(ctx, "briggs", [(p.freq, bu, bw, n)], robust=0.0)
        ts.append(time.perf_counter() - t0)
    out["briggs_weighting_ms"] = min(ts) * 1e3
    bw.upload(p.weight)
    for name, k in (("pillbox_1x1", GriddingKernel.pillbox()),
                    ("pillbox_7x7", GriddingKernel.build("pillbox", p.M, p.N, p.deltax, p.deltay,
                                                         m=7, n=7))):
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            gu, _, _ = grid_visibilities_dev(ctx, k, p.freq, bu, bv, bw, n)
            ts.append(time.perf_counter() - t0)
        out[name] = {"ms": min(ts) * 1e3, "rows_per_s": n / min(ts), "gridded_rows": gu.shape[0],
                     "note": "includes copying the gridded rows to the host"}
    if threads > 0:
        sys.path.insert(0, str(ROOT / "oracle"))
        os.environ["OMP_NUM_THREADS"] = str(threads)
        import oracle

        cpu_s, k = gridding_comparator(p, oracle, threads)
        out["cpu_comparator_7x7"] = {"seconds": cpu_s, "threads": threads,
                                     "sample": f"{k} of {n} rows, extrapolated linearly"}
        out["speedup_7x7"] = cpu_s / (out["pillbox_7x7"]["ms"] * 1e-3)
    # the objective step on the gridded rows (default 1x1 pillbox)
    s, ng = add_gridded_dev(ctx, GriddingKernel.pillbox(), p.freq, bu, bv, bw, n)
    terms = [("Chi2", -1, 0.0), ("Entropy", 0, 0.001), ("TotalSquaredVariation", 1, 0.0)]
    obj = Objective(ctx, fw_config(terms, lambdas=(0.01, 1e-3)))
    d_img = ctx.image(bench_image(p))
    xi = ctx.zeros(2 * p.M * p.N)
    for _ in range(warmup):
        obj.value(d_img)
        obj.gradient(d_img, xi, 1)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        obj.value(d_img)
        obj.gradient(d_img, xi, 1)
    ctx.sync()
    el = time.perf_counter() - t0
    out["gridded_objective"] = {"rows": ng, "evals_per_s": steps / el,
                                "ms_per_step": el / steps * 1e3, "adjoint": adjoint}
    obj.close()
    ctx.close()
    return out


def generate_pairs(seeds, iterations: int, which: int):
    """Training pairs from the plain solver for `seeds`, with a progress line
    after each pair."""
    import numpy as np
    from gpuvmem_amd import warmstart as W

    t0 = time.perf_counter()
    seeds = list(seeds)
    chunks = []
    for c0 in range(len(seeds)):
        chunks.append(W.make_pairs(seeds[c0:c0 + 1], iterations=iterations, which=which))
        print(f"# pairs {c0 + 1} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    return W.Pairs(np.concatenate([c.dirty for c in chunks]),
                   np.concatenate([c.model for c in chunks]), seeds,
                   time.perf_counter() - t0)


def bench_warm_start(n_train: int, n_test: int, steps: int, iterations: int, levels: int,
                     base: int, which: int = 0, batch: int = 8, load=None,
                     eval_lbfgs: int = 0, save_model=None) -> dict:
    """The learned warm start on observations of workload `which` (0: 256^2 /
    36,000 visibilities; 1: 1024^2 / 1,764,000; 4: 512^2 / 1,029,000): training pairs from the plain
    solver (CG-FRPRMN to its tolerance; generated here, or loaded from the
    files `load` that --ws-save runs wrote), U-Net training on the device
    kernels (the trained network written to `save_model` when given), then
    the plain solver from the default image and from the prediction on
    held-out observations."""
    import numpy as np
    from gpuvmem_amd import warmstart as W
    from gpuvmem_amd.device import Context

    t0 = time.perf_counter()
    if load:
        parts = [np.load(f) for f in load]
        pairs = W.Pairs(np.concatenate([q["dirty"] for q in parts]),
                        np.concatenate([q["model"] for q in parts]),
                        [int(x) for q in parts for x in q["seeds"]])
        # generated by separate processes side by side: the longest one
        t_pairs = max(float(q["seconds"]) for q in parts)
        n_train = len(pairs.dirty)
        print(f"# {n_train} pairs from {len(parts)} files", file=sys.stderr, flush=True)
    else:
        pairs = generate_pairs(range(1000, 1000 + n_train), iterations, which)
        t_pairs = pairs.seconds
    p = W.observation(1000, which)
    ctx = Context(p, add_data=False)
    ws = W.WarmStart(ctx, levels=levels, base=base, batch=batch, npix=p.M)
    t_train = ws.train(pairs, steps=steps, log_every=25, progress=lambda k, loss: print(
        f"# step {k} loss {loss:.3e} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr,
        flush=True))
    print(f"# trained in {t_train:.1f} s, loss {ws.history[-1][1]:.3e}", file=sys.stderr, flush=True)
    if save_model:
        ws.save(save_model, meta={
            "workload": which, "image": [p.M, p.N], "visibilities": p.nvis,
            "pairs": n_train, "pair_seeds": [min(pairs.seeds), max(pairs.seeds)],
            "pair_solver": f"CG-FRPRMN on chi2+entropy, at most {iterations} iterations, "
                           "bf16x3 adjoint",
            "levels": levels, "base_channels": base, "parameters": ws.net.nparams,
            "adam_steps": steps, "batch": batch, "learning_rate": 1e-3,
            "final_loss": ws.history[-1][1], "train_seconds": t_train,
            "pairs_seconds": t_pairs})
        print(f"# network written to {save_model}", file=sys.stderr, flush=True)
    if n_test <= 0:
        out = {"trained": True, "pairs": n_train, "pairs_seconds": t_pairs,
               "train_seconds": t_train, "train_loss": ws.history, "model": save_model}
        ws.close()
        ctx.close()
        return out
    ev = eval_settings(iterations, eval_lbfgs)
    r = W.evaluate(ws, range(5000, 5000 + n_test), plain=True, which=which, **ev,
                   progress=lambda row: print(
                       f"# held-out {row['seed']}: iterations {row['iters_default']} / "
                       f"{row['iters_warm']}, seconds {row['seconds_default']:.2f} / "
                       f"{row['seconds_warm']:.2f}", file=sys.stderr, flush=True))
    rows = r.pop("rows")
    out = {
        "metric": "learned warm start: time to the objective both starts reach, default / warm",
        "value": r["median_time_ratio"], "unit": "x (median over held-out observations)",
        "higher_is_better": True, "n_gpus": 1, "data": "synthetic (seeded generator)",
        "config": {"workload": f"{p.M}x{p.N} image, {p.nvis} visibilities, pairs from "
                               f"chi2+entropy, CG-FRPRMN to its tolerance (at most {iterations} "
                               "iterations); evaluated with " +
                               (f"L-BFGS m=10 on chi2+entropy+TSV (at most {eval_lbfgs} "
                                "iterations)" if eval_lbfgs else "the same solver"),
                   "unet": {"levels": levels, "base_channels": base, "parameters": ws.net.nparams,
                            "weights": "bf16 (fp32 master, fp32 accumulation)" if ws.net.bf16
                            else "fp32", "batch": batch},
                   "train_pairs": n_train, "adam_steps": steps, "held_out": n_test,
                   "pair_files": len(load) if load else 0},
        "pairs_seconds": t_pairs, "train_seconds": t_train,
        "train_loss": ws.history, **r,
        "rows": [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in row.items()}
                 for row in rows],
    }
    ws.close()
    ctx.close()
    return out


def eval_settings(iterations: int, eval_lbfgs: int) -> dict:
    """The held-out solves: the pair solver (CG-FRPRMN on chi2 + entropy, at
    most `iterations`), or with eval_lbfgs the time-to-image objective and
    optimiser (L-BFGS m = 10 on chi2 + entropy + TSV, at most eval_lbfgs)."""
    if eval_lbfgs:
        return dict(iterations=eval_lbfgs, optimizer="CG-LBFGS",
                    terms=[("Chi2", -1, 0.0), ("Entropy", 0, 0.001),
                           ("TotalSquaredVariation", 1, 0.0)], lambdas=(0.01, 1e-3))
    return dict(iterations=iterations)


def bench_warm_eval(model: str, seeds, which: int, iterations: int, eval_lbfgs: int = 0) -> dict:
    """A network written by --ws-save-model on the held-out observations
    `seeds` of workload `which`: the plain solver from the default image and
    from the prediction, iterations and seconds to the objective both reach
    (dirty image and inference counted for the warm start), one record per
    observation."""
    from gpuvmem_amd import warmstart as W
    from gpuvmem_amd.device import Context

    seeds = list(seeds)
    p = W.observation(seeds[0], which)
    ctx = Context(p, add_data=False)
    ws = W.WarmStart.load(ctx, model, batch=1, npix=p.M)
    r = W.evaluate(ws, seeds, which=which, **eval_settings(iterations, eval_lbfgs),
                   progress=lambda row: print(
                       f"# held-out {row['seed']}: iterations {row['iters_default']} / "
                       f"{row['iters_warm']}, seconds {row['seconds_default']:.2f} / "
                       f"{row['seconds_warm']:.2f}", file=sys.stderr, flush=True))
    r["network"] = ws.meta | {"file": Path(model).name}
    r["rows"] = [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in row.items()}
                 for row in r["rows"]]
    ws.close()
    ctx.close()
    return r


def bench_unet_step(steps: int, warmup: int, levels: int = 6, base: int = 32, batch: int = 8,
                    npix: int = 512) -> dict:
    """Training steps of the warm-start U-Net on synthetic batches: forward,
    MSE loss, backward and one Adam update per step, on the project's own
    kernels (FP32-MFMA implicit-GEMM convolutions, BF16 weights)."""
    from gpuvmem_amd.device import Context, DeviceBuffer
    from gpuvmem_amd.problem import make_problem
    from gpuvmem_amd.warmstart import UNet

    ctx = Context(make_problem(0, npix=64), add_data=False)
    net = UNet(ctx, levels=levels, base=base, batch=batch, H=npix, W=npix, seed=1, bf16=True)
    rng = np.random.default_rng(2)
    net.set_input(rng.standard_normal((batch, 1, npix, npix)).astype(np.float32))
    target = DeviceBuffer.from_array(rng.standard_normal((batch, 1, npix, npix)).astype(np.float32))

    def step():
        net.forward()
        net.loss(target)
        net.backward()
        net.adam(lr=1e-4)

    for _ in range(max(1, warmup)):
        step()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.sync()
    el = (time.perf_counter() - t0) / steps
    out = {"metric": "warm-start network training step (forward + MSE + backward + Adam)",
           "value": el * 1e3, "unit": "ms per step", "higher_is_better": False, "n_gpus": 1,
           "steps": steps, "data": "synthetic",
           "config": {"levels": levels, "base_channels": base, "batch": batch,
                      "image": [npix, npix], "parameters": net.nparams,
                      "weights": "bf16 (fp32 master)",
                      "convolutions": "implicit GEMM on v_mfma_f32_32x32x2_f32"}}
    net.close()
    ctx.close()
    return out


def profiled_traffic(mode: str, nvis: int, split_vis: int):
    """HBM bytes per launch of the adjoint GEMM from the committed rocprofv3
    counter summaries (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of
    known size), when one matches this adjoint form, visibility count and
    range length (the number of partial images the GEMM writes follows from
    the last two)."""
    pat = re.compile(r"^traffic_bytes_per_launch mode=(\S+) nvis=(\d+) split_vis=(\d+) "
                     r"bytes=(\d+)", re.M)
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for m in pat.finditer(f.read_text()):
            if m.group(1) == mode and int(m.group(2)) == nvis and int(m.group(3)) == split_vis:
                return int(m.group(4)), f.name
    return None, None


def copy_calibration(ctx, mib: int) -> None:
    """A plain device-to-device copy of known size (the library's copy
    kernel), outside the timed region, so counter runs can scale FETCH_SIZE /
    WRITE_SIZE against it."""
    from gpuvmem_amd import native
    from gpuvmem_amd.device import DeviceBuffer

    n = (mib << 20) // 4
    src, dst = DeviceBuffer(4 * n), DeviceBuffer(4 * n)
    native.check(ctx.lib.gvm_vec_zero(ctx.ctx, src.ptr, n), "zero")
    for _ in range(3):
        native.check(ctx.lib.gvm_vec_copy(ctx.ctx, dst.ptr, src.ptr, n), "copy")
    ctx.sync()
    src.free()
    dst.free()
    print(f"# copy calibration: copy_kernel of {4 * n} bytes read and {4 * n} bytes written, "
          "3 launches", file=sys.stderr, flush=True)


def _slice_problem(p, k):
    import copy

    q = copy.copy(p)
    q.slices = None
    q.uvw = p.uvw[:k]
    q.vo = p.vo[:k]
    q.weight = p.weight[:k]
    return q


def time_to_image(p, iterations: int, forms=("f32", "bf16x3"), warm_model=None) -> dict:
    """L-BFGS (memory 10) from the default initial image with each adjoint
    form, up to `iterations` iterations (the reference's optimiser may stop
    earlier when its line search makes no progress).  The target is the
    objective value every form reached (the largest of the final values),
    and each form is timed to its first iteration at or below it.  Each
    context does one untimed evaluation first (workspace allocation).  Every
    form runs with the same settings: evaluation graphs replayed and the
    adjoint range length autotuned, so the ratio measures the adjoint's
    precision form alone.

    With `warm_model` (a network written by WarmStart.save), the last form
    runs once more from the learned warm start: the dirty image, the
    network's prediction and the fallback check (the default image when the
    prediction's objective is not lower) are timed with the run.  Its target
    is the value all runs reached."""
    from gpuvmem_amd.device import Context, Objective, fw_config

    terms = [("Chi2", -1, 0.0), ("Entropy", 0, 0.001), ("TotalSquaredVariation", 1, 0.0)]
    out, runs, images = {}, {}, {}
    legs = [(m, False) for m in forms] + ([(forms[-1], True)] if warm_model else [])
    for mode, warm_leg in legs:
        name = mode + "_warm" if warm_leg else mode
        extras = mode != "direct"
        ctx = Context(p, adjoint=mode)
        obj = Objective(ctx, fw_config(terms, lambdas=(0.01, 1e-3)))
        obj.set_graphs(extras)
        warm = ctx.image(p.initial_image())
        xi = ctx.zeros(2 * p.M * p.N)
        obj.value(warm)
        tune = ctx.autotune() if extras else None
        obj.gradient(warm, xi, 0)
        rec = {}
        if warm_leg:
            from gpuvmem_amd import warmstart as W

            ws = W.WarmStart.load(ctx, warm_model, batch=1)
            W.start_image(ws, ctx, obj, p)[0].free()  # untimed first run of the network
            ctx.sync()
            t0 = time.perf_counter()
            d, used, f_start = W.start_image(ws, ctx, obj, p)
            ctx.sync()
            t_start = time.perf_counter() - t0
            rec = {"prediction_used": used, "start_objective": f_start,
                   "start_seconds": t_start, "network": ws.meta | {"file": Path(warm_model).name}}
            ws.close()
        else:
            ctx.sync()
            d = ctx.image(p.initial_image())
            f_start, t_start = obj.value(d), 0.0
        res, traj, secs = obj.optimize_timed(d, "CG-LBFGS", iterations, lbfgs_k=10)
        obj.value(d)
        chi2 = obj.term_values()[0]
        images[name] = d.download(np.float32, (2, p.N, p.M))
        # the starting point first, then one entry per iteration
        runs[name] = ([f_start] + list(traj), [t_start] + [t_start + s for s in secs])
        out[name] = {
            "iterations_run": res.iterations,
            "seconds_run": t_start + res.seconds,
            "seconds_per_iteration": res.seconds / max(1, res.iterations),
            "final_objective": res.final_value,
            "final_chi2": chi2,
            "function_calls": res.function_calls,
            "gradient_calls": res.gradient_calls,
            "graphs": extras,
            "graph_replays": obj.graph_launches(),
            "split_vis": ctx.split_vis() if mode != "direct" else None,
            "autotune": tune,
            **rec,
        }
        obj.close()
        ctx.close()
    names = list(runs)
    # the split ratio's target: the value the default-start forms reached
    target = max(runs[m][0][-1] for m in forms)
    target_all = max(runs[m][0][-1] for m in names)
    for name in names:
        traj, secs = runs[name]
        for key, tg in (("", target), ("_all", target_all)):
            if key == "" and name not in forms:
                continue
            hit = next(k for k, v in enumerate(traj) if v <= tg)
            out[name]["seconds_to_target" + key] = secs[hit]
            out[name]["iterations_to_target" + key] = hit
        print(f"# time-to-image {name}: {out[name]}", file=sys.stderr, flush=True)
    ref = images[forms[0]]
    for name in names[1:]:
        out[name]["image_rel_l2_vs_" + forms[0]] = float(
            np.linalg.norm(images[name] - ref) / np.linalg.norm(ref))
    return {"target_objective": target, "target_objective_all": target_all, "forms": out}


def tti_observations(p, workload: int, iterations: int, n_obs: int, warm_model=None) -> dict:
    """time_to_image on n_obs observations (new sky and noise seeds, same
    array): the line search is chaotic, so where each form stops, and the
    ratio at the common objective, vary from one observation to the next.
    Returns the median ratio of plain FP32 seconds to split-precision
    seconds with the per-observation records, and with `warm_model` the
    median ratio of plain FP32 seconds to split-precision seconds from the
    learned warm start, both to the value all runs reached."""
    obs = []
    for k in range(max(1, n_obs)):
        pk = p if k == 0 else make_problem_of(workload, seed_sky=p.sim.seed_sky + 1000 * k,
                                              seed_noise=p.sim.seed_noise + 1000 * k)
        r = time_to_image(pk, iterations, warm_model=warm_model)
        f32, split = r["forms"]["f32"], r["forms"]["bf16x3"]
        r["ratio"] = (f32["seconds_to_target"] / split["seconds_to_target"]
                      if split["seconds_to_target"] else None)
        r["per_iteration_ratio"] = (f32["seconds_per_iteration"] /
                                    split["seconds_per_iteration"])
        r["seed_sky"] = int(pk.sim.seed_sky)
        msg = f"# observation {k}: ratio {r['ratio']:.3f}, per iteration {r['per_iteration_ratio']:.3f}"
        if warm_model:
            w = r["forms"]["bf16x3_warm"]
            r["ratio_warm"] = f32["seconds_to_target_all"] / w["seconds_to_target_all"]
            r["ratio_split_all"] = f32["seconds_to_target_all"] / split["seconds_to_target_all"]
            msg += f", warm start {r['ratio_warm']:.3f}"
        print(msg, file=sys.stderr, flush=True)
        obs.append(r)
    ratios = [o["ratio"] for o in obs if o["ratio"]]
    out = {
        "value": float(np.median(ratios)) if ratios else None,
        "unit": "x (plain FP32 seconds / split-precision seconds, median over observations)",
        "config": {"workload": f"{p.M}x{p.N} image, {p.nvis} visibilities, chi2+entropy+TSV, "
                               f"L-BFGS m=10, at most {iterations} iterations per form",
                   "observations": len(obs),
                   "settings": "both forms: evaluation graphs, autotuned adjoint range length"},
        "median_per_iteration_ratio": float(np.median([o["per_iteration_ratio"] for o in obs])),
        "observations": obs,
    }
    if warm_model:
        ws = [o["forms"]["bf16x3_warm"] for o in obs]
        out["warm_start"] = {
            "value": float(np.median([o["ratio_warm"] for o in obs])),
            "unit": "x (plain FP32 seconds / split-precision + graphs + learned warm start "
                    "seconds, to the objective all runs reached; median over observations)",
            "median_ratio_split_same_target": float(np.median([o["ratio_split_all"]
                                                               for o in obs])),
            "median_iterations": {
                "f32": float(np.median([o["forms"]["f32"]["iterations_to_target_all"]
                                        for o in obs])),
                "bf16x3": float(np.median([o["forms"]["bf16x3"]["iterations_to_target_all"]
                                           for o in obs])),
                "bf16x3_warm": float(np.median([w["iterations_to_target_all"] for w in ws]))},
            "prediction_used": sum(bool(w["prediction_used"]) for w in ws),
            "median_start_seconds": float(np.median([w["start_seconds"] for w in ws])),
            "network": ws[0]["network"],
        }
    return out


def make_problem_of(workload: int, **overrides):
    from gpuvmem_amd.problem import make_problem

    return make_problem(workload, **overrides)


def precision_study(p, pixels: int = 64, split_vis: int = 0) -> dict:
    """Each adjoint form against FP64: chi2 against the oracle's FP64
    forward pass, the gradient against FP64 direct sums on a pixel sample
    inside the mask, and the split forms against the FP32 and direct forms
    over the whole image."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    from gpuvmem_amd.device import Context

    img = bench_image(p)
    orc = oracle.Oracle(p, separable=True)
    ref_chi2, vr, _ = orc.chi2(img.astype(np.float64))
    rng = np.random.default_rng(7)
    inside = np.argwhere(p.noise < p.noise_cut)
    pix = inside[rng.choice(len(inside), pixels, replace=False)]
    gref = orc.dchi2_pixels(vr, pix)
    out, full = {}, {}
    for mode in ("direct", "f32", "f16x3", "bf16x3"):
        ctx = Context(p, adjoint=mode)
        if split_vis and mode != "direct":
            ctx.set_split_vis(split_vis)
        d = ctx.image(img)
        v = ctx.chi2(d)
        res = ctx.zeros(2 * p.M * p.N)
        ctx.dchi2(d, res)  # allocates the workspace
        ctx.sync()
        ctx.reset_timing()
        res = ctx.zeros(2 * p.M * p.N)
        ctx.dchi2(d, res)
        ctx.sync()
        ms, _ = ctx.adjoint_time_ms()
        g = res.download(np.float32, (2, p.N, p.M))[0].astype(np.float64)
        ctx.close()
        gs = np.array([g[i, j] for i, j in pix])
        full[mode] = g
        out[mode] = {"chi2_rel_err": abs(v - ref_chi2) / abs(ref_chi2),
                     "grad_rel_l2_vs_fp64_pixels": float(np.linalg.norm(gs - gref) / np.linalg.norm(gref)),
                     "adjoint_ms": ms}
        print(f"# precision {mode}: {out[mode]}", file=sys.stderr, flush=True)
    for mode in ("f32", "f16x3", "bf16x3"):
        for base in ("direct", "f32"):
            if mode != base:
                out[mode][f"grad_rel_l2_vs_{base}_image"] = float(
                    np.linalg.norm(full[mode] - full[base]) / np.linalg.norm(full[base]))
    return {"pixels": pixels, "forms": out}


def replay_image_study(p, iterations: int, forms=("bf16x3", "f16x3", "f32/6144")) -> dict:
    """The image after `iterations` CG-FRPRMN iterations (chi2 + entropy +
    TSV from the default image) with the plain FP32 adjoint, its applied step
    lengths recorded, then the same path in each other form with those steps
    replayed: what separates the final images is then the adjoint's
    arithmetic, not the line search's branch decisions.  Reports each form's
    final image (inside the mask) and objective values against the FP32
    run's.  A form "mode/L" runs with visibility ranges of L: "f32/6144" is
    the FP32 form with only its accumulation order changed (at 2048^2 the
    default ranges are 3072 visibilities, and the workspace limit keeps them
    from getting shorter), which shows how far the path carries a
    rounding-level difference in the gradient."""
    from gpuvmem_amd.device import Context, Objective, fw_config

    terms = [("Chi2", -1, 0.0), ("Entropy", 0, 0.001), ("TotalSquaredVariation", 1, 0.0)]
    mask = p.noise < p.noise_cut
    runs = {}
    for mode in ("f32",) + tuple(forms):
        form, _, rng = mode.partition("/")
        ctx = Context(p, adjoint=form)
        if rng:
            ctx.set_split_vis(int(rng))
        obj = Objective(ctx, fw_config(terms, lambdas=(0.01, 1e-3)))
        if mode != "f32":
            obj.set_line_search_replay(runs["f32"]["steps"])
        d = ctx.image(p.initial_image())
        n = iterations if mode == "f32" else len(runs["f32"]["steps"])
        t0 = time.perf_counter()
        res, traj = obj.optimize(d, "CG-FRPRMN", n)
        sec = time.perf_counter() - t0
        full = d.download(np.float32, (2, p.N, p.M))
        img0 = full[0][mask]
        runs[mode] = {"res": res, "traj": np.array(traj), "steps": obj.line_search_steps(),
                      "img": img0.astype(np.float64), "seconds": sec, "full": full,
                      # pixels held at the positivity floor of the image update
                      "floor": img0 <= np.float32(p.minpix)}
        print(f"# replay study {mode}: {res.iterations} iterations, f {res.final_value:.10e}, "
              f"{sec:.1f} s", file=sys.stderr, flush=True)
        obj.close()
        ctx.close()
    ref = runs["f32"]
    # the chi2 gradient at the end of the FP32 path, every form against FP32 over the mask
    # and against FP64 direct sums on sampled pixels (as at the start of the path)
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    final = ref["full"]
    orc = oracle.Oracle(p, separable=True)
    _, vr, _ = orc.chi2(final.astype(np.float64))
    inside = np.argwhere(mask)
    pix = inside[np.random.default_rng(7).choice(len(inside), 64, replace=False)]
    gref = orc.dchi2_pixels(vr, pix)
    grads = {}
    for mode in ("f32",) + tuple(forms):
        form, _, rng = mode.partition("/")
        ctx = Context(p, adjoint=form)
        if rng:
            ctx.set_split_vis(int(rng))
        d = ctx.image(final)
        ctx.chi2(d)
        res = ctx.zeros(2 * p.M * p.N)
        ctx.dchi2(d, res)
        ctx.sync()
        grads[mode] = res.download(np.float32, (2, p.N, p.M))[0].astype(np.float64)
        ctx.close()
    grad_final = {}
    for mode, g in grads.items():
        gs = np.array([g[i, j] for i, j in pix])
        grad_final[mode] = {"rel_l2_vs_fp64_pixels": float(np.linalg.norm(gs - gref) /
                                                           np.linalg.norm(gref))}
        if mode != "f32":
            grad_final[mode]["rel_l2_vs_f32_mask"] = float(
                np.linalg.norm((g - grads["f32"])[mask]) / np.linalg.norm(grads["f32"][mask]))
    out_grad = {"gradient_at_final_image": grad_final}
    out = {"reference": {"form": "f32", "iterations": ref["res"].iterations,
                         "final_objective": ref["res"].final_value, "seconds": ref["seconds"],
                         "pixels_in_mask": int(mask.sum()),
                         "pixels_at_floor": int(ref["floor"].sum())}}
    for mode in forms:
        r = runs[mode]
        k = min(len(r["traj"]), len(ref["traj"]))
        out[mode] = {
            "iterations": r["res"].iterations,
            "final_image_rel_l2_vs_f32": float(np.linalg.norm(r["img"] - ref["img"]) /
                                               np.linalg.norm(ref["img"])),
            "final_objective_rel_vs_f32": abs(r["res"].final_value - ref["res"].final_value) /
                                          abs(ref["res"].final_value),
            "max_path_value_rel_vs_f32": float(np.max(np.abs(r["traj"][:k] - ref["traj"][:k]) /
                                                      np.abs(ref["traj"][:k]))),
            "pixels_at_floor": int(r["floor"].sum()),
            "pixels_floor_differs": int((r["floor"] != ref["floor"]).sum()),
            "final_image_rel_l2_off_floor": float(
                np.linalg.norm((r["img"] - ref["img"])[~(r["floor"] | ref["floor"])]) /
                np.linalg.norm(ref["img"][~(r["floor"] | ref["floor"])])),
            "seconds_replayed": r["seconds"]}
    out.update(out_grad)
    return out


def gradient_difference_study(p, iterations: int,
                              forms=("f32", "f32/6144", "bf16x3", "f16x3")) -> dict:
    """How each form's gradient error changes between two consecutive
    images of an optimisation path, against the change of the gradient
    itself: the FP32 form's CG path (chi2 + entropy + TSV) gives image B after
    `iterations` steps and, its first steps replayed, image A one step
    before.  With the direct form as the reference g (its error is an order
    below the others'), each form's error d = g_form - g is taken at A and
    B on the mask; Polak-Ribiere's gamma and L-BFGS's y see
    |d(B) - d(A)| / |g(B) - g(A)|."""
    from gpuvmem_amd.device import Context, Objective, fw_config

    terms = [("Chi2", -1, 0.0), ("Entropy", 0, 0.001), ("TotalSquaredVariation", 1, 0.0)]
    mask = p.noise < p.noise_cut
    imgs, steps = {}, None
    for name, n in (("B", iterations), ("A", iterations - 1)):
        ctx = Context(p, adjoint="f32")
        obj = Objective(ctx, fw_config(terms, lambdas=(0.01, 1e-3)))
        if steps is not None:
            obj.set_line_search_replay(steps[:n])
        d = ctx.image(p.initial_image())
        obj.optimize(d, "CG-FRPRMN", n)
        if steps is None:
            steps = obj.line_search_steps()
        imgs[name] = d.download(np.float32, (2, p.N, p.M))
        obj.close()
        ctx.close()
    grads = {}
    for mode in ("direct",) + tuple(forms):
        form, _, rng = mode.partition("/")
        ctx = Context(p, adjoint=form)
        if rng:
            ctx.set_split_vis(int(rng))
        for name, img in imgs.items():
            d = ctx.image(img)
            ctx.chi2(d)
            res = ctx.zeros(2 * p.M * p.N)
            ctx.dchi2(d, res)
            ctx.sync()
            grads[mode, name] = res.download(np.float32, (2, p.N, p.M))[0][mask].astype(np.float64)
        ctx.close()
        print(f"# difference study: {mode} done", file=sys.stderr, flush=True)
    gA, gB = grads["direct", "A"], grads["direct", "B"]
    dg = np.linalg.norm(gB - gA)
    out = {"iterations": iterations, "steps": len(steps),
           "gradient_change_rel": float(dg / np.linalg.norm(gB))}
    for mode in forms:
        eA, eB = grads[mode, "A"] - gA, grads[mode, "B"] - gB
        out[mode] = {"error_rel_at_B": float(np.linalg.norm(eB) / np.linalg.norm(gB)),
                     # a systematic offset: the error's mean over its rms, and its
                     # correlation with the gradient itself
                     "error_mean_over_rms_at_B": float(eB.mean() / np.sqrt(np.mean(eB * eB))),
                     "error_correlation_with_gradient_at_B": float(
                         np.dot(eB, gB) / (np.linalg.norm(eB) * np.linalg.norm(gB))),
                     "error_change_over_gradient_change": float(np.linalg.norm(eB - eA) / dg),
                     "error_correlation_A_B": float(np.dot(eA, eB) /
                                                    (np.linalg.norm(eA) * np.linalg.norm(eB)))}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--adjoint", default="bf16x3", choices=list(PEAK_TFLOPS))
    ap.add_argument("--workload", type=int, default=1,
                    help="1: 1024^2/1.76M vis, 0: 256^2/36k, 2: 2048^2/4.35M vis, 3: "
                         "multi-frequency 512^2 I_nu_0 + alpha, 4 spw x 64 channels, ~61M vis")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--tti-observations", type=int, default=5,
                    help="time-to-image: observations (sky and noise seeds) to take the median over")
    ap.add_argument("--tti-iterations", type=int, default=100,
                    help="time-to-image leg of the default run: L-BFGS iteration cap per form")
    ap.add_argument("--tti-warm-model", metavar="FILE", default=str(WARM_MODEL),
                    help="time to image: also from the learned warm start of this network "
                         "(written by --ws-save-model; 'none' to skip)")
    ap.add_argument("--no-tti", action="store_true",
                    help="default run: leave out the time-to-image and plain-FP32 legs")
    ap.add_argument("--time-to-image", type=int, default=0, metavar="ITERS",
                    help="instead of evals/s: L-BFGS time-to-image of the split-precision "
                         "adjoint against the plain FP32 one at equal objective value")
    ap.add_argument("--study-iterations", type=int, default=0, metavar="ITERS",
                    help="precision study: also the image after ITERS CG-FRPRMN iterations, "
                         "FP32 path replayed in the split forms")
    ap.add_argument("--split-vis", type=int, default=0, metavar="L",
                    help="precision study: visibility range length of the GEMM forms "
                         "(default: the context's)")
    ap.add_argument("--diff-study", type=int, default=0, metavar="ITERS",
                    help="precision study: each form's gradient error between the last two "
                         "images of an ITERS-step FP32 CG path, against the gradient's change")
    ap.add_argument("--study-forms", default="bf16x3,f16x3,f32/6144",
                    help="precision study: the forms that replay the FP32 path (mode or mode/L)")
    ap.add_argument("--precision-study", action="store_true",
                    help="instead of evals/s: each adjoint form against FP64 (use with "
                         "--workload 2: 2048^2, 4.35M vis)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the adjoint's range length by timing before the run")
    ap.add_argument("--graphs", action="store_true",
                    help="replay captured HIP graphs of the value and gradient evaluations")
    ap.add_argument("--warm-start", action="store_true",
                    help="train the U-Net warm start and measure its gain (256^2 observations)")
    ap.add_argument("--ws-pairs", type=int, default=96, help="warm start: training pairs")
    ap.add_argument("--ws-steps", type=int, default=400, help="warm start: Adam steps")
    ap.add_argument("--ws-test", type=int, default=12, help="warm start: held-out observations")
    ap.add_argument("--ws-no-graphs", action="store_true",
                    help="warm start: the solver launches every kernel (no evaluation graphs)")
    ap.add_argument("--ws-workload", type=int, default=0, choices=(0, 1, 4),
                    help="warm start: observations of workload 0 (256^2, 36,000 vis), 1 "
                         "(1024^2, 1,764,000 vis) or 4 (512^2, 1,029,000 vis)")
    ap.add_argument("--ws-levels", type=int, default=None,
                    help="warm start: U-Net levels (default 4 at 256^2, 6 otherwise)")
    ap.add_argument("--ws-base", type=int, default=None,
                    help="warm start: U-Net base channels (default 16 at 256^2 and 1024^2, 32 "
                         "at 512^2)")
    ap.add_argument("--ws-iterations", type=int, default=500,
                    help="warm start: the solver's iteration cap (pairs and evaluation)")
    ap.add_argument("--ws-eval-lbfgs", type=int, default=0, metavar="ITERS",
                    help="warm start: evaluate with L-BFGS (m=10) on chi2 + entropy + TSV, the "
                         "time-to-image objective, for at most ITERS iterations")
    ap.add_argument("--ws-batch", type=int, default=8, help="warm start: training batch")
    ap.add_argument("--ws-save", metavar="FILE",
                    help="warm start: only generate the pairs of --ws-seeds and save them (.npz)")
    ap.add_argument("--ws-seeds", metavar="A:B", default=None,
                    help="warm start: sky seeds [A, B) of the pairs that --ws-save generates")
    ap.add_argument("--ws-load", nargs="+", metavar="FILE",
                    help="warm start: train on the pairs of these files instead of generating")
    ap.add_argument("--ws-save-model", metavar="FILE", default=None,
                    help="warm start: write the trained network to FILE (.npz); "
                         "with --ws-test 0, train and write only")
    ap.add_argument("--ws-eval-model", metavar="FILE", default=None,
                    help="warm start: only evaluate the network in FILE on the held-out "
                         "observations --ws-seeds A:B")
    ap.add_argument("--unet-step", action="store_true",
                    help="time training steps (forward, MSE, backward, Adam) of the warm-start "
                         "network at its size: 6 levels, 32 base channels, 512^2, batch 8, BF16 "
                         "weights")
    ap.add_argument("--gridding", action="store_true",
                    help="time the device weighting and gridder and the gridded objective step")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the run, copy a buffer of this size (for counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU split over visibilities is a design note only)", file=sys.stderr)
        return 2

    from gpuvmem_amd.device import Context, Objective, fw_config
    from gpuvmem_amd.problem import make_mfs_problem, make_problem

    if args.unet_step:
        print(json.dumps(bench_unet_step(args.steps, args.warmup)), flush=True)
        return 0
    if args.warm_start:
        if args.ws_no_graphs:
            from gpuvmem_amd import warmstart

            warmstart.GRAPHS = False
        if args.ws_eval_model:
            a, b = (int(x) for x in (args.ws_seeds or "5000:%d" % (5000 + args.ws_test)).split(":"))
            r = bench_warm_eval(args.ws_eval_model, range(a, b), args.ws_workload,
                                args.ws_iterations, args.ws_eval_lbfgs)
            print(json.dumps(r), flush=True)
            return 0
        if args.ws_save:
            a, b = (int(x) for x in (args.ws_seeds or "1000:%d" % (1000 + args.ws_pairs)).split(":"))
            pr = generate_pairs(range(a, b), args.ws_iterations, args.ws_workload)
            with open(args.ws_save, "wb") as f:
                np.savez(f, dirty=pr.dirty, model=pr.model, seeds=np.array(pr.seeds),
                         seconds=pr.seconds)
            print(json.dumps({"saved": args.ws_save, "pairs": len(pr.dirty),
                              "seconds": pr.seconds}), flush=True)
            return 0
        r = bench_warm_start(n_train=args.ws_pairs, n_test=args.ws_test, steps=args.ws_steps,
                             iterations=args.ws_iterations,
                             levels=args.ws_levels or (4 if args.ws_workload == 0 else 6),
                             base=args.ws_base or (32 if args.ws_workload == 4 else 16),
                             which=args.ws_workload, batch=args.ws_batch, load=args.ws_load,
                             eval_lbfgs=args.ws_eval_lbfgs, save_model=args.ws_save_model)
        print(json.dumps(r), flush=True)
        return 0
    t0 = time.perf_counter()
    p = make_mfs_problem() if args.workload == 3 else make_problem(args.workload)
    print(f"# workload {args.workload}: {p.M}x{p.N}, {len(p.all_slices())} slices, "
          f"{p.nvis_total} visibilities "
          f"(generated in {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    if args.precision_study:
        r = precision_study(p, split_vis=args.split_vis)
        if args.diff_study > 0:
            r["gradient_difference"] = gradient_difference_study(p, args.diff_study)
        if args.study_iterations > 0:
            r["replayed_path"] = replay_image_study(p, args.study_iterations,
                                                    tuple(args.study_forms.split(",")))
        print(json.dumps({"metric": f"adjoint precision against FP64 ({p.M}^2, {p.nvis_total} vis)",
                          "n_gpus": 1, "data": "synthetic", **r}), flush=True)
        return 0
    if args.gridding:
        threads = 0 if args.no_cpu_baseline else host_cpus()["threads"]
        r = bench_gridding(p, args.steps, args.warmup, threads, args.adjoint)
        print(json.dumps({"metric": f"gridded-visibility set-up and step ({p.M}^2, "
                                    f"{p.nvis_total} vis)", "n_gpus": 1, "data": "synthetic",
                          **r}), flush=True)
        return 0
    if args.time_to_image > 0:
        r = tti_observations(p, args.workload, args.time_to_image, args.tti_observations,
                             warm_model=tti_warm_model(args, p))
        print(json.dumps({
            "metric": "time-to-image vs plain FP32 at equal objective (1024^2, 1.76M vis)",
            "higher_is_better": True, "n_gpus": 1, "data": "synthetic", **r}), flush=True)
        return 0
    ctx = Context(p, adjoint=args.adjoint)
    terms = [("Chi2", -1, 0.0), ("Entropy", 0, 0.001), ("TotalSquaredVariation", 1, 0.0)]
    obj = Objective(ctx, fw_config(terms, lambdas=(0.01, 1e-3)))
    obj.set_graphs(args.graphs)
    d_img = ctx.image(bench_image(p))
    xi = ctx.zeros(2 * p.M * p.N)
    tune = None
    if args.autotune:
        obj.value(d_img)  # residuals for the timed adjoints
        tune = ctx.autotune()
        print(f"# autotune: {tune}", file=sys.stderr, flush=True)

    def step():
        obj.value(d_img)
        obj.gradient(d_img, xi, 1)

    # with graphs, the third evaluation of a kind is the first replay
    for _ in range(max(args.warmup, 3) if args.graphs else args.warmup):
        step()
    ctx.sync()
    ctx.reset_timing()
    replays0 = obj.graph_launches()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    elapsed = time.perf_counter() - t0
    replays = obj.graph_launches() - replays0
    if args.graphs:
        # the adjoint's launch timing is taken outside graphs: a few eager steps
        obj.set_graphs(False)
        ctx.reset_timing()
        for _ in range(3):
            step()
        ctx.sync()
    adj_ms, adj_n = ctx.adjoint_time_ms()
    print(f"# {args.steps} steps in {elapsed:.3f} s; adjoint kernel {adj_ms:.3f} ms avg over "
          f"{adj_n} launches", file=sys.stderr, flush=True)

    nvt = p.nvis_total
    nsl = len(p.all_slices())
    traffic, traffic_src = profiled_traffic(args.adjoint, nvt, ctx.split_vis())
    # the adjoint computes the sub-tiles that hold a pixel on the mask: the work
    # of a launch is 4 flop per (computed pixel, visibility)
    px = ctx.adjoint_computed_pixels()
    flops = 4.0 * px * nvt
    achieved = flops / (adj_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[args.adjoint]
    line = {
        "metric": ("obj+grad evals/s (1024^2, 1.76M vis)" if args.workload == 1 else
                   f"obj+grad evals/s ({p.M}^2, {nsl} slices, {nvt} vis)"),
        "value": args.steps / elapsed,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.adjoint],
        "data": ("synthetic (seeded generator: 50-antenna array, 200 Gaussian components, 1 mJy "
                 "noise)" if args.workload == 1 else "synthetic (seeded generator)"),
        "config": {
            "workload": (f"{p.M}x{p.N} image, {nvt} visibilities, {nsl} channel(s), "
                         "chi2+entropy+TSV, CG/L-BFGS objective step"),
            "image": [p.M, p.N],
            "visibilities": nvt,
            "slices": nsl,
            "adjoint": args.adjoint,
            "graphs": bool(args.graphs),
            "graph_replays": replays,
            "split_vis": ctx.split_vis(),
            "autotune": tune,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic,
            "traffic_unit": "bytes/launch (HBM, rocprofv3 FETCH_SIZE+WRITE_SIZE)",
            "traffic_source": traffic_src,
            "kernel": f"adjoint_gemm_kernel ({args.adjoint})",
            "kernel_ms": adj_ms,
            "algorithmic_flop_per_launch": flops,
            "computed_pixels": px,
            "image_pixels": p.M * p.N,
            "full_image_flop_per_launch": 4.0 * p.M * p.N * nvt,
        },
    }
    line["value_evaluation_us"] = value_latency(obj, d_img, ctx)
    if args.calibrate_copy_mib > 0:
        copy_calibration(ctx, args.calibrate_copy_mib)
    obj.close()
    ctx.close()
    if not args.no_tti and args.workload == 1 and args.adjoint != "f32":
        # the plain FP32 form's evals/s on the same step, then time to image
        line["plain_fp32"] = plain_fp32_step(p, args.steps, args.warmup, args.graphs)
        print(f"# plain FP32 step: {line['plain_fp32']}", file=sys.stderr, flush=True)
        tti = tti_observations(p, args.workload, args.tti_iterations, args.tti_observations,
                               warm_model=tti_warm_model(args, p))
        line["time_to_image"] = compact_tti(tti)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(p, host_cpus(), args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


def tti_warm_model(args, p):
    """The network for the warm-started time-to-image leg: --tti-warm-model
    when the file exists and was trained at this image size, else None."""
    f = args.tti_warm_model
    if not f or f.lower() == "none" or not Path(f).is_file():
        return None
    with np.load(f, allow_pickle=False) as z:
        if int(z["npix"]) != p.M or p.M != p.N:
            print(f"# {f}: trained at {int(z['npix'])}^2, not {p.M}^2; no warm-start leg",
                  file=sys.stderr, flush=True)
            return None
    return f


def value_latency(obj, d_img, ctx, n: int = 200) -> dict:
    """Wall time of one value-only evaluation (ObjectiveFunction::calcFunction,
    what each line-search probe costs: forward chi2 and prior values, read
    back to the host), with evaluation graphs replayed and without."""
    out = {}
    for graphs in (True, False):
        obj.set_graphs(graphs)
        for _ in range(3):
            obj.value(d_img)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(n):
            obj.value(d_img)
        out["graphs" if graphs else "eager"] = (time.perf_counter() - t0) / n * 1e6
    obj.set_graphs(False)
    out["evaluations"] = n
    return out


def plain_fp32_step(p, steps: int, warmup: int, graphs: bool) -> dict:
    """The benchmark's step with the plain FP32 MFMA adjoint."""
    from gpuvmem_amd.device import Context, Objective, fw_config

    ctx = Context(p, adjoint="f32")
    terms = [("Chi2", -1, 0.0), ("Entropy", 0, 0.001), ("TotalSquaredVariation", 1, 0.0)]
    obj = Objective(ctx, fw_config(terms, lambdas=(0.01, 1e-3)))
    obj.set_graphs(graphs)
    d_img = ctx.image(bench_image(p))
    xi = ctx.zeros(2 * p.M * p.N)
    for _ in range(max(warmup, 3) if graphs else max(warmup, 1)):
        obj.value(d_img)
        obj.gradient(d_img, xi, 1)
    ctx.sync()
    ctx.reset_timing()
    t0 = time.perf_counter()
    for _ in range(steps):
        obj.value(d_img)
        obj.gradient(d_img, xi, 1)
    ctx.sync()
    el = time.perf_counter() - t0
    ms, n = ctx.adjoint_time_ms()
    split_vis = ctx.split_vis()
    obj.close()
    ctx.close()
    return {"evals_per_s": steps / el, "ms_per_step": el / steps * 1e3, "steps": steps,
            "adjoint_kernel_ms": ms if n else None, "split_vis": split_vis, "graphs": bool(graphs)}


def compact_tti(r: dict) -> dict:
    """The time-to-image summary with one short record per observation."""
    obs = []
    for o in r["observations"]:
        rec = {"seed_sky": o["seed_sky"], "ratio": o["ratio"],
               "per_iteration_ratio": o["per_iteration_ratio"],
               "target_objective": o["target_objective"]}
        if "ratio_warm" in o:
            rec["ratio_warm"] = o["ratio_warm"]
            rec["target_objective_all"] = o["target_objective_all"]
        for m, f in o["forms"].items():
            rec[m] = {k: f[k] for k in ("seconds_to_target", "iterations_to_target",
                                        "seconds_to_target_all", "iterations_to_target_all",
                                        "iterations_run", "seconds_run", "final_objective",
                                        "final_chi2", "split_vis", "prediction_used",
                                        "start_objective", "start_seconds",
                                        "image_rel_l2_vs_f32") if k in f}
        obs.append(rec)
    keep = ("value", "unit", "config", "median_per_iteration_ratio", "warm_start")
    return {k: r[k] for k in keep if k in r} | {"observations": obs}


if __name__ == "__main__":
    sys.exit(main())
