"""rama_amd — MI355X-native primal-dual multicut solver (RAMA's hot path).

The solver runs as hand-written HIP kernels for gfx950 inside
librama_hip.so; Python reaches it through the C ABI (`rama_amd._capi`) or the
pybind11 module `rama_py`, which keeps RAMA's Python interface.
"""
from . import _capi  # noqa: F401
from ._capi import RamaError, options, solve, solve_device, objective  # noqa: F401

__all__ = ["RamaError", "options", "solve", "solve_device", "objective"]
