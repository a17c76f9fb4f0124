"""MI355X (gfx950) implementation of gpuvmem's maximum-entropy imaging
objective and gradient.

Native layout:
  include/gvm_hip.h          C ABI of the HIP library (lib/libgvm_hip.so)
  include/gvm/*.h            C++ framework: Fi, ObjectiveFunction, Optimizer,
                             factories, set-up, synthetic data (lib/libgvm.so)
Python modules are thin ctypes harnesses over those libraries.
"""
from .native import GvmError, load_fw, load_hip  # noqa: F401

__all__ = ["GvmError", "load_fw", "load_hip"]
