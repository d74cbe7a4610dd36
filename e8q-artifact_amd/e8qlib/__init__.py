"""E8Q and ME8Q IVF indexes on AMD MI355X (gfx950).

Drop-in for the reference package's `e8qlib.E8qIndex` and
`e8qlib.Me8qIndex`: same constructors,
build/search/save/load signatures and outputs, with build (E8 encode) and
search running as HIP kernels on the GPU.  There is no CPU fallback: if the
native extension or its device library is missing, importing this package
fails.

PyTorch, when installed, is imported first: it ships its own HIP runtime
(libamdhip64.so.7), and loading it before libe8q_gpu.so makes the process
use that one runtime for both, so device pointers from torch tensors are
valid in search_device().
"""
import os

try:  # one HIP runtime per process (see above)
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional
    pass

_here = os.path.dirname(os.path.abspath(__file__))
if not os.path.exists(os.path.join(_here, "libe8q_gpu.so")):
    raise ImportError("e8qlib: libe8q_gpu.so is not built (run __graft_entry__.build())")

from ._e8qlib import (E8qIndex, Me8qIndex, Server, search_context_info, set_batched_shape,  # noqa: E402
                      set_me8q_grids, set_search_graphs)

__all__ = ["E8qIndex", "Me8qIndex", "Server", "set_search_graphs", "set_batched_shape", "set_me8q_grids",
           "search_context_info"]
