"""gpu4gst: optimal group Steiner trees (TrimCDP / D-TrimCDP) on MI355X.

The solver is libgst.so (hand-written gfx950 HIP kernels behind a C ABI,
declared in include/gst.h); this package is a thin ctypes layer over it plus
the torch.autograd.Function of the differentiable cost.
"""
from .solver import (GSTError, GSTGraph, SolveResult, BatchResult, dp_gpu, dp_gpu_hop,
                     STATUS_OK, STATUS_INFEASIBLE, STATUS_RANGE, STATUS_NO_TREE)

__all__ = ["GSTError", "GSTGraph", "SolveResult", "BatchResult", "dp_gpu", "dp_gpu_hop",
           "STATUS_OK", "STATUS_INFEASIBLE", "STATUS_RANGE", "STATUS_NO_TREE"]
