"""The committed warm-start network of the 1024^2 workload: readable without
pickle, shaped as its record says, finite, and trained on observations
disjoint from the ones the benchmark's time-to-image leg runs on."""
import json
from pathlib import Path

import numpy as np
import pytest

from gpuvmem_amd.problem import sim_config
from gpuvmem_amd.warmstart import unet_shapes

NET = Path(__file__).resolve().parents[1] / "gpuvmem_amd" / "data" / "warm_start_1024sq.npz"


@pytest.fixture(scope="module")
def net():
    if not NET.is_file():
        pytest.skip("no committed network")
    with np.load(NET, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_shapes_match_the_record(net):
    levels, base = int(net["levels"]), int(net["base"])
    shapes = unet_shapes(levels, base)
    n_w = sum(co * ci * k * k for _, co, ci, k in shapes)
    n_b = sum(co for _, co, _, _ in shapes)
    assert bool(net["bf16"])
    assert net["weights_bf16"].dtype == np.uint16 and net["weights_bf16"].shape == (n_w,)
    assert net["biases"].dtype == np.float32 and net["biases"].shape == (n_b,)
    meta = json.loads(str(net["meta"]))
    assert meta["parameters"] == n_w + n_b
    assert (meta["levels"], meta["base_channels"]) == (levels, base)
    assert meta["workload"] == 1 and int(net["npix"]) == 1024 == meta["image"][0]
    assert int(net["npix"]) % (1 << (levels - 1)) == 0


def test_values_are_finite(net):
    w = (net["weights_bf16"].astype(np.uint32) << 16).view(np.float32)
    assert np.isfinite(w).all() and np.isfinite(net["biases"]).all()
    assert np.isfinite(float(net["calib"])) and float(net["calib"]) > 0
    # trained, not the initialisation: no all-zero layer
    assert np.abs(w).max() > 0


def test_training_observations_are_not_benchmarked(net):
    """Pairs come from observation(seed) (sky seed = seed); the time-to-image
    leg runs the workload's own sky seed and four more 1000 apart."""
    lo, hi = json.loads(str(net["meta"]))["pair_seeds"]
    s0 = sim_config(1).seed_sky
    bench = [s0 + 1000 * k for k in range(5)]
    assert not any(lo <= s <= hi for s in bench)
