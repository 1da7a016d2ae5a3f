"""MI355X-native Vertex Block Descent (gfx950 HIP kernels behind a C ABI).

Importing the package loads libvbd_hip.so; there is no CPU fallback for the solver.
"""
from . import meshgen  # noqa: F401
from ._capi import VBDError  # noqa: F401
from .solver import (DeviceMesh, TetMesh, VBD_GPU_Solver, VBDParams, VBDsolver,  # noqa: F401
                     mesh_from_generator, prepare_host, upload)

__all__ = ["meshgen", "VBDError", "DeviceMesh", "TetMesh", "VBD_GPU_Solver", "VBDParams", "VBDsolver",
           "mesh_from_generator", "prepare_host", "upload"]
