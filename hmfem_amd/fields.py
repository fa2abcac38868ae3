"""Seeded permeability fields (host C library libhmfem_fields.so).

Thin ctypes layer over include/hmfem_fields.h. Arrays are float64, length
nx*ny*nz, index s = i + j*nx + k*nx*ny.
"""
import ctypes

import numpy as np

from ._lib import load_fields

_c_int, _c_dbl, _c_u64, _c_vp = ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_void_p


def _lib():
    lib = load_fields()
    if not getattr(lib, "_typed", False):
        lib.hmfem_field_homogeneous.argtypes = [_c_int, _c_int, _c_int, _c_dbl, _c_vp]
        lib.hmfem_field_lognormal.argtypes = [_c_int, _c_int, _c_int, _c_dbl, _c_dbl, _c_dbl, _c_u64, _c_vp]
        lib.hmfem_field_conduits.argtypes = [_c_int, _c_int, _c_int, _c_int, _c_u64, _c_vp]
        lib.hmfem_field_instance_seed.argtypes = [_c_u64, _c_int]
        lib.hmfem_field_instance_seed.restype = _c_u64
        lib.hmfem_field_uniform.argtypes = [ctypes.POINTER(_c_u64)]
        lib.hmfem_field_uniform.restype = _c_dbl
        lib._typed = True
    return lib


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with code %d" % (what, rc))


def homogeneous(nx, ny, nz, k):
    K = np.empty(nx * ny * nz, np.float64)
    _check(_lib().hmfem_field_homogeneous(nx, ny, nz, k, K.ctypes.data), "hmfem_field_homogeneous")
    return K


def lognormal(nx, ny, nz, mean_ln, sigma, corr_len, seed):
    K = np.empty(nx * ny * nz, np.float64)
    _check(_lib().hmfem_field_lognormal(nx, ny, nz, mean_ln, sigma, corr_len, seed, K.ctypes.data),
           "hmfem_field_lognormal")
    return K


def conduits(K, nx, ny, nz, count, seed):
    """Multiply K in place along `count` seeded conduits; returns K."""
    assert K.dtype == np.float64 and K.flags.c_contiguous and K.size == nx * ny * nz
    _check(_lib().hmfem_field_conduits(nx, ny, nz, count, seed, K.ctypes.data), "hmfem_field_conduits")
    return K


def instance_seed(seed, instance):
    return int(_lib().hmfem_field_instance_seed(seed, instance))


class Stream:
    """Per-instance parameter draws (uniform in [0,1)) from the same PRNG."""

    def __init__(self, seed, instance):
        self._s = _c_u64(_lib().hmfem_field_instance_seed(seed, instance))

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * _lib().hmfem_field_uniform(ctypes.byref(self._s))

    def choice(self, items):
        return items[min(int(self.uniform() * len(items)), len(items) - 1)]
