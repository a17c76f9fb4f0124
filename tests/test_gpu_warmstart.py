"""What the trained warm-start network gains over the method without it, on
held-out observations: a much lower starting objective and fewer optimiser
iterations to the objective value both starts reach (a small training run:
24 solver pairs, 150 Adam steps)."""
import numpy as np
import pytest

from gpuvmem_amd import warmstart as W
from gpuvmem_amd.device import Context

pytestmark = pytest.mark.gpu


def test_warm_start_gains_over_the_default_image():
    print("case: warm start 24 pairs / 150 steps / 4 held-out", flush=True)
    pairs = W.make_pairs(range(2000, 2024), iterations=200)
    print(f"pairs in {pairs.seconds:.1f} s", flush=True)
    p = W.observation(2000)
    ctx = Context(p, add_data=False)
    ws = W.WarmStart(ctx, levels=4, base=8, batch=8, npix=p.M, seed=1)
    t = ws.train(pairs, steps=150)
    print(f"trained in {t:.1f} s; loss {ws.history[0][1]:.3e} -> {ws.history[-1][1]:.3e}", flush=True)
    assert ws.history[-1][1] < 0.5 * ws.history[0][1]
    r = W.evaluate(ws, range(6000, 6004), iterations=300)
    for row in r["rows"]:
        print({k: row[k] for k in ("seed", "f0_default", "f0_warm", "iters_default", "iters_warm")})
    # the prediction starts far below the constant image on every observation
    assert all(row["f0_warm"] < 0.5 * row["f0_default"] for row in r["rows"])
    # and reaches the common objective in fewer iterations (median over seeds)
    print(f"median iteration ratio {r['median_iter_ratio']:.2f}")
    assert r["median_iter_ratio"] > 1.0
    ws.close()
    ctx.close()


@pytest.mark.parametrize("bf16", [True, False])
def test_saved_network_predicts_bit_for_bit(tmp_path, bf16):
    """WarmStart.save / WarmStart.load: the loaded network's prediction
    equals the saved one's bit for bit (BF16 weight bits and FP32 biases, or
    FP32 everything), after a few Adam steps moved the FP32 masters off
    BF16 values."""
    print(f"case: save / load round trip, bf16={bf16}, 3 levels x 4 channels at 256^2", flush=True)
    p = W.observation(2100)
    ctx = Context(p)
    dirty = W.dirty_image(ctx, p)
    ws = W.WarmStart(ctx, levels=3, base=4, batch=2, npix=p.M, seed=3, bf16=bf16)
    pairs = W.Pairs(np.stack([dirty, 0.5 * dirty]), np.stack([np.abs(dirty), 0.3 * np.abs(dirty)]))
    ws.train(pairs, steps=3)
    a = ws.predict(dirty[None])
    f = tmp_path / "net.npz"
    ws.save(f, meta={"case": "round trip"})
    ws2 = W.WarmStart.load(ctx, f, batch=2)
    b = ws2.predict(dirty[None])
    assert ws2.meta == {"case": "round trip"}
    assert ws2.calib == ws.calib
    assert np.array_equal(a, b)
    ws.close()
    ws2.close()
    ctx.close()


def test_committed_network_starts_below_the_default_image():
    """The committed 1024^2 network on the benchmark's observation (not one
    it was trained on): the prediction is used (the fallback rule keeps the
    default image otherwise) and the objective of the time-to-image runs
    starts well below the default image's."""
    from pathlib import Path

    from gpuvmem_amd.device import Objective, fw_config
    from gpuvmem_amd.problem import make_problem

    f = Path(__file__).resolve().parents[1] / "gpuvmem_amd" / "data" / "warm_start_1024sq.npz"
    if not f.is_file():
        pytest.skip("no committed network")
    print("case: committed warm-start network, workload 1 observation", flush=True)
    p = make_problem(1)
    ctx = Context(p)
    obj = Objective(ctx, fw_config([("Chi2", -1, 0.0), ("Entropy", 0, 0.001),
                                    ("TotalSquaredVariation", 1, 0.0)], lambdas=(0.01, 1e-3)))
    ws = W.WarmStart.load(ctx, f)
    d, used, f_start = W.start_image(ws, ctx, obj, p)
    d0 = ctx.image(p.initial_image())
    f_default = obj.value(d0)
    print(f"start objective {f_start:.6e} against {f_default:.6e} at the default image", flush=True)
    assert used
    assert f_start < 0.5 * f_default
    d.free()
    d0.free()
    ws.close()
    obj.close()
    ctx.close()
