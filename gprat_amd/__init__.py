"""gprat_amd — MI355X (gfx950) backend for GPRat's exact Gaussian-process path.

`import gprat_amd as gprat` gives the reference's Python API (GP, GP_data,
AdamParams, compute_*_tiles, start_hpx, ...) backed by hand-written HIP
kernels; `gprat_amd.capi` binds the C ABI directly. The native extension is
required: importing fails loudly if it has not been built.
"""
import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))
LIBRARY_PATH = _os.path.join(_HERE, "libgprat_amd.so")

try:
    from .gprat import *  # noqa: F401,F403
    from .gprat import GP, GP_data, AdamParams, SEKParams, Model, HIP_GPU  # noqa: F401
except ImportError as _e:  # pragma: no cover - exercised only on an unbuilt tree
    raise ImportError(
        "gprat_amd: the native extension is not built (run __graft_entry__.build() or "
        "python gprat_amd/build.py)"
    ) from _e
