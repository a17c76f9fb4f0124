"""The parts of bench.py that run without a GPU: the HBM traffic figures it
takes from the committed rocprofv3 counter summaries, and the refusal of more
than one GPU."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402

NVIS = 1_764_000
# the GEMM writes one FP32 1024 x 1024 partial image per visibility range, on the
# mask only: ranges of 6144 visibilities (the default), their count padded to a
# multiple of 8 (288)
SPLIT = 6144
RANGES = -(-(-(-NVIS // SPLIT)) // 8) * 8


@pytest.fixture(scope="module")
def masked_pixels():
    from gpuvmem_amd.problem import make_problem
    p = make_problem(1)
    assert p.nvis == NVIS
    return int((p.noise < p.noise_cut).sum())


@pytest.mark.parametrize("mode", ["bf16x3", "f16x3"])
def test_traffic_summary_for_each_mfma_form(mode, masked_pixels):
    traffic, source = bench.profiled_traffic(mode, NVIS, SPLIT)
    assert source is not None, f"no counter summary for {mode} at {NVIS} visibilities"
    assert (ROOT / "profiles" / source).exists()
    # at least the partial tiles it writes on the mask, at most 15 % more than that
    # (whole 32-byte sectors at the mask's edge, and the reads)
    partial_bytes = RANGES * masked_pixels * 4
    assert partial_bytes <= traffic <= 1.15 * partial_bytes, (traffic, partial_bytes)
    text = (ROOT / "profiles" / source).read_text()
    assert f"--adjoint {mode}" in text or mode == "bf16x3"


def test_no_summary_for_an_unprofiled_size_or_range_length():
    assert bench.profiled_traffic("bf16x3", NVIS + 1, SPLIT) == (None, None)
    # another range length writes another number of partial images
    assert bench.profiled_traffic("bf16x3", NVIS, 1536) == (None, None)


def test_host_cpu_facts():
    hc = bench.host_cpus()
    assert hc["nproc"] >= 1 and 1 <= hc["threads"] <= hc["nproc"]
    if hc["cgroup_cpu_quota"] is not None:
        assert hc["threads"] <= max(1, int(hc["cgroup_cpu_quota"]))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
