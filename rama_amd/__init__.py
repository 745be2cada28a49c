"""rama_amd — MI355X-native primal-dual multicut solver (RAMA's hot path).

The solver runs as hand-written HIP kernels for gfx950 inside
librama_hip.so; Python reaches it through the C ABI (`rama_amd._capi`) or the
pybind11 module `rama_py`, which keeps RAMA's Python interface.
"""
# One HIP runtime per process: torch ships its own libamdhip64 with the same
# soname as the system one, and torch only works on its own.  Loading torch
# first makes librama_hip.so bind to that runtime too.
try:  # pragma: no cover - depends on the environment
    import torch  # noqa: F401
except ImportError:
    pass

from . import _capi  # noqa: F401,E402
from ._capi import RamaError, Server, objective, options, solve, solve_batch, solve_device  # noqa: F401,E402

__all__ = ["RamaError", "options", "solve", "solve_device", "solve_batch", "Server", "objective"]
