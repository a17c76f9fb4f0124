"""Builds the native libraries in-tree.

* ``lib/libgvm_hip.so`` -- HIP kernels + the C ABI of ``include/gvm_hip.h``,
  compiled by hipcc for gfx950 only.
* ``lib/libgvm.so``     -- the C++ framework (Fi / ObjectiveFunction /
  optimisers / factories), set-up helpers and the synthetic-data generator,
  compiled by g++ and linked against libgvm_hip.so (rpath $ORIGIN).

Run ``python -m gpuvmem_amd.build`` or call :func:`build`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
INCLUDE = ROOT / "include"

HIP_SOURCES = ["gvm_ctx.hip", "gvm_forward.hip", "gvm_adjoint.hip", "gvm_ops.hip", "gvm_setup.hip",
               "gvm_grid.hip", "gvm_nn.hip"]
HOST_SOURCES = ["framework.cpp", "optimizers.cpp", "setup.cpp", "fw_capi.cpp", "simdata.cpp",
                "fitsio.cpp", "ckernels.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Packed-FP32 VALU forms are disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in res.stdout.splitlines() if "packed-fp32-ops' is not a recognized" not in line
    )
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build_hip() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libgvm_hip.so"
    tmp = LIB / "libgvm_hip.so.tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", f"-I{INCLUDE}"]
    cmd += [str(CSRC / s) for s in HIP_SOURCES]
    cmd += ["-o", str(tmp)]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_host() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libgvm.so"
    tmp = LIB / "libgvm.so.tmp"
    # no multiply-add contraction: host float arithmetic in source order
    cmd = [CXX, "-std=c++17", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared", "-Wall",
           f"-I{INCLUDE}"]
    cmd += [str(CSRC / "host" / s) for s in HOST_SOURCES]
    cmd += [f"-L{LIB}", "-lgvm_hip", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build() -> None:
    if not shutil.which(HIPCC) and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    build_hip()
    build_host()


if __name__ == "__main__":
    build()
    print("built", LIB / "libgvm_hip.so", LIB / "libgvm.so")
