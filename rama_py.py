"""`import rama_py` — RAMA's Python module (src/rama_py.cu:76-125), served by
rama_amd's pybind11 extension over the HIP library."""
from rama_amd.rama_py import (  # noqa: F401
    multicut_solver_options,
    rama_cuda,
    rama_cuda_gpu_pointers,
    rama_torch,
    read_multicut_file,
)
