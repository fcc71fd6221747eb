"""pptk_amd: pptk's k-d tree and normal estimation on AMD Instinct MI355X.

Drop-in for the two native entry points of pptk (pptk/__init__.py:5-6):

* ``pptk_amd.kdtree`` -- the ``kdtree`` extension module with ``_build`` and
  ``_query``;
* ``pptk_amd.estimate_normals`` -- the ``estimate_normals`` function.

Both run on hand-written gfx950 HIP kernels in ``lib/libpptk_amd.so``; there
is no CPU fallback. Importing raises ImportError when the native modules have
not been built (``python -m pptk_amd._build``).
"""
from ._runtime import preload_hip_runtime

preload_hip_runtime()

from .kdtree import kdtree  # noqa: E402
from .processing.estimate_normals.estimate_normals import estimate_normals  # noqa: E402

__all__ = ["kdtree", "estimate_normals"]

from . import capi  # noqa: E402,F401  (ctypes binding of the C ABI)
