"""MI355X (gfx950) byte-pattern replace: bytes.replace semantics on HBM-resident data.

The compute path is libcuda_replace.so (hand-written HIP kernels + C++ runtime,
built in-tree by ``python -m cuda_replace_amd.build``).  This package exposes it
through the reference wrapper's API.
"""
from .wrapper import (DEFAULT_LIB, REQUIRED_SYMBOLS, BatchEngine, CudaReplaceLib,  # noqa: F401
                      Server, Session, Ticket, gpu_replace_streaming, replace_unified)

__all__ = ["CudaReplaceLib", "Session", "BatchEngine", "Server", "Ticket", "replace_unified",
           "gpu_replace_streaming",
           "DEFAULT_LIB", "REQUIRED_SYMBOLS"]
