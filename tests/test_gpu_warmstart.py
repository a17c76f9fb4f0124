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
