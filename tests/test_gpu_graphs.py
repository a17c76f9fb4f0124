"""HIP graphs of enqueued evaluations: a captured value + gradient replayed
several times, with eager work and new image contents in between, gives the
eager results bit for bit; stale graphs and waiting calls are refused; the
framework's graph cache leaves optimiser trajectories unchanged."""
import ctypes as C

import numpy as np
import pytest

from gpuvmem_amd import native
from gpuvmem_amd.device import Context, Objective, fw_config
from gpuvmem_amd.problem import small_problem

pytestmark = pytest.mark.gpu

TERMS = [("Chi2", -1, 0.0), ("Entropy", 0, 0.001), ("TotalSquaredVariation", 1, 0.0)]
LAMBDAS = (0.01, 1e-3)


def _img(p, seed):
    rng = np.random.default_rng(seed)
    img = p.initial_image()
    img[0] += (rng.random((p.N, p.M)) * 20.0).astype(np.float32)
    return img


def _eager(ctx, d, res, n):
    ctx.chi2_enqueue(d, 0)
    st = ctx.lib.gvm_entropy_value_enqueue(ctx.ctx, d.ptr, 0, 0.001, -1.0, 1)
    native.check(st, "entropy enqueue")
    ctx.values_fetch_enqueue(2)
    vals = ctx.values_host(2)
    native.check(ctx.lib.gvm_vec_zero(ctx.ctx, res.ptr, n), "zero")
    ctx.dchi2(d, res)
    ctx.sync()
    return vals, res.download(np.float32)


@pytest.mark.parametrize("mode", ["bf16x3", "f32"])
def test_captured_evaluation_replays_bit_for_bit(mode):
    p = small_problem(npix=128, n_antennas=8, seed=21)
    print(f"case: graph replay {mode} {p.M}^2 nvis={p.nvis}", flush=True)
    ctx = Context(p, adjoint=mode)
    n = 2 * p.M * p.N
    d = ctx.image(_img(p, 0))
    res = ctx.zeros(n)
    _eager(ctx, d, res, n)  # sizes the adjoint workspace
    ctx.capture_begin()
    ctx.chi2_enqueue(d, 0)
    native.check(ctx.lib.gvm_entropy_value_enqueue(ctx.ctx, d.ptr, 0, 0.001, -1.0, 1), "ent")
    ctx.values_fetch_enqueue(2)
    native.check(ctx.lib.gvm_vec_zero(ctx.ctx, res.ptr, n), "zero")
    ctx.dchi2(d, res)
    g = ctx.capture_end()
    for seed in (1, 2, 3):
        img = _img(p, seed)
        # eager evaluation of this image first, then the graph on the same buffers
        d.upload(img)
        ref_vals, ref_grad = _eager(ctx, d, res, n)
        d.upload(img)
        g.launch()
        vals = ctx.values_host(2)
        grad = res.download(np.float32)
        print(f"seed {seed}: chi2 {vals[0]:.6e}")
        assert vals == ref_vals
        assert np.array_equal(grad, ref_grad)
    # a new adjoint form moves the context to a new epoch: the graph is refused
    ctx.set_adjoint("f16x3" if mode != "f16x3" else "bf16x3")
    assert g.status_of_launch() == 4  # GVM_ERR_STATE
    g.destroy()
    ctx.close()


def test_waiting_calls_are_refused_while_capturing():
    p = small_problem(npix=64, n_antennas=6, seed=22)
    print(f"case: capture refusals {p.M}^2 nvis={p.nvis}", flush=True)
    ctx = Context(p)
    d = ctx.image(_img(p, 0))
    res = ctx.zeros(2 * p.M * p.N)
    ctx.chi2(d)
    ctx.capture_begin()
    ctx.chi2_enqueue(d, 0)
    # the adjoint workspace is not sized yet: refused instead of waiting
    assert ctx.lib.gvm_dchi2(ctx.ctx, d.ptr, res.ptr, 0, 0) != 0
    assert ctx.lib.gvm_sync(ctx.ctx) != 0
    out = (C.c_double * 1)()
    assert ctx.lib.gvm_values_read(ctx.ctx, 1, out) != 0
    g = ctx.capture_end()
    g.destroy()
    # the context still works
    v = ctx.chi2(d)
    assert np.isfinite(v)
    ctx.close()


@pytest.mark.parametrize("optimizer", ["CG-FRPRMN", "CG-LBFGS"])
def test_objective_graph_cache_keeps_trajectories(optimizer):
    p = small_problem(npix=128, n_antennas=8, seed=23)
    print(f"case: graph cache {optimizer} {p.M}^2 nvis={p.nvis}", flush=True)
    out = []
    for graphs in (False, True):
        ctx = Context(p)
        obj = Objective(ctx, fw_config(TERMS, lambdas=LAMBDAS))
        obj.set_graphs(graphs)
        d = ctx.image(p.initial_image())
        res, traj = obj.optimize(d, optimizer=optimizer, iterations=6,
                                 lbfgs_k=10 if optimizer == "CG-LBFGS" else 0)
        ctx.sync()
        img = d.download(np.float32)
        launches = obj.graph_launches()
        out.append((traj, img, launches, res.function_calls, res.gradient_calls))
        obj.close()
        ctx.close()
    (t0, i0, l0, nv0, ng0), (t1, i1, l1, nv1, ng1) = out
    print(f"evaluations {nv0}+{ng0}; graph replays {l1}")
    assert l0 == 0 and l1 > 0
    assert t0 == t1
    assert np.array_equal(i0, i1)
