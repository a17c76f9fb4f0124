"""The C++ mirror of gpuvmem's plugin interfaces used the way gpuvmem's own
main program composes them (tests/cpp/compose_main.cpp: globals, terms and
optimiser created by factory id, src/main.cu:147-203): the program compiles
and links against include/ and the two libraries with nothing else, and on the
GPU its optimiser trajectory equals, bit for bit, the one the flat C entry
(gvm_fw_*) gives for the same objective."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "gpuvmem_amd" / "lib"
SRC = ROOT / "tests" / "cpp" / "compose_main.cpp"


def _build(tmp_path) -> Path:
    exe = tmp_path / "compose"
    cmd = ["g++", "-std=c++17", "-O2", f"-I{ROOT / 'include'}", str(SRC), "-o", str(exe),
           f"-L{LIB}", "-lgvm", "-lgvm_hip", f"-Wl,-rpath,{LIB}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_composition_compiles_and_links(tmp_path):
    assert (LIB / "libgvm.so").exists() and (LIB / "libgvm_hip.so").exists()
    assert _build(tmp_path).exists()


def _write_input(p, path: Path) -> None:
    with open(path, "wb") as f:
        f.write(bytes(p.geometry()))
        f.write(np.ascontiguousarray(p.noise, np.float32).tobytes())
        slices = p.all_slices()
        f.write(np.int32(len(slices)).tobytes())
        for freq, uvw, vo, w in slices:
            f.write(np.float32(freq).tobytes())
            f.write(np.int64(uvw.shape[0]).tobytes())
            f.write(np.ascontiguousarray(uvw, np.float64).tobytes())
            f.write(np.ascontiguousarray(vo, np.float32).tobytes())
            f.write(np.ascontiguousarray(w, np.float32).tobytes())


@pytest.mark.gpu
@pytest.mark.parametrize("optimizer", ["CG-FRPRMN", "CG-LBFGS"])
def test_composition_matches_flat_entry(tmp_path, optimizer):
    from gpuvmem_amd.device import Context, Objective, fw_config
    from gpuvmem_amd.problem import small_problem

    p = small_problem(npix=64, n_antennas=8, seed=17)
    print(f"case: C++ composition {optimizer}, {p.M}^2 nvis={p.nvis}", flush=True)
    exe = _build(tmp_path)
    inp = tmp_path / "input.bin"
    _write_input(p, inp)
    iters = 8
    r = subprocess.run([str(exe), str(inp), optimizer, str(iters)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    line = next(l for l in r.stdout.splitlines() if l.startswith("trajectory"))
    traj_cpp = np.array([float(v) for v in line.split()[1:]])

    ctx = Context(p)
    obj = Objective(ctx, fw_config([("Chi2", -1, 0.0), ("Entropy", 0, 0.001)], lambdas=(0.01,)))
    d = ctx.image(p.initial_image())
    _, traj = obj.optimize(d, optimizer, iters, lbfgs_k=10)
    obj.close()
    ctx.close()
    traj = np.array(traj)
    print(f"C++: {traj_cpp}\nflat: {traj}")
    assert len(traj_cpp) >= 2 and np.array_equal(traj_cpp, traj)
    # the mapping filled the reference's way (float steps only): the same
    # first value, and then values that differ only by the float rounding of
    # each accepted step (relative 6e-8 on the step)
    r = subprocess.run([str(exe), str(inp), optimizer, str(iters), "float-steps"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    line = next(l for l in r.stdout.splitlines() if l.startswith("trajectory"))
    traj_f = np.array([float(v) for v in line.split()[1:]])
    n = min(len(traj_f), len(traj))
    rel = np.abs(traj_f[:n] - traj[:n]) / np.abs(traj[:n])
    print(f"float steps: {traj_f}; max rel {rel.max():.2e}")
    assert n >= 2 and rel.max() < 1e-3
