"""One HIP runtime per process.

torch ships its own copy of the HIP runtime (torch/lib/libamdhip64.so, with
the same soname as ROCm's libamdhip64.so.7). Whichever copy a process loads
first satisfies later requests for that soname; if this package's library
were loaded first it would bring ROCm's copy, torch would then load its own
as a second runtime and fail to start. Loading torch's copy first (when
torch is installed) keeps a single runtime whatever the import order, so
device pointers and streams created by torch are valid in this library.
Nothing here imports torch.
"""
import ctypes
import importlib.util
import os

_done = False


def preload_hip_runtime():
    global _done
    if _done:
        return
    _done = True
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if spec is None or not spec.submodule_search_locations:
        return
    for d in spec.submodule_search_locations:
        path = os.path.join(d, "lib", "libamdhip64.so")
        if os.path.exists(path):
            try:
                ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
            except OSError:
                pass
            return
