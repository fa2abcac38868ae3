"""Loaders for the package's native libraries (built in-tree by build.py)."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
GPU_LIB = os.path.join(HERE, "libhmfem_gpu.so")
FIELDS_LIB = os.path.join(HERE, "libhmfem_fields.so")

_cache = {}


def _load(path, what):
    if path not in _cache:
        if not os.path.exists(path):
            raise RuntimeError("%s is not built (%s missing); run `python -c 'import __graft_entry__ as g; g.build()'`"
                               % (what, path))
        _cache[path] = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    return _cache[path]


def load_fields():
    return _load(FIELDS_LIB, "the field generator library")


def load_gpu():
    return _load(GPU_LIB, "the HIP solver library")
