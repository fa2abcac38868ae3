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


# ------------------------------------------------------------------ block right-hand sides

def face_index_sizes(nx, ny, nz):
    """Lengths of the six face value arrays (faces l, r, u, d, t, b): face l/r
    is indexed k*ny + j, u/d k*nx + i, t/b j*nx + i (set_source_term.c)."""
    return (ny * nz, ny * nz, nx * nz, nx * nz, nx * ny, nx * ny)


def dirichlet_pair_faces(nx, ny, nz, axis, forward):
    """Face values of an all-Dirichlet box with p = 1 on one face of `axis`
    (0 x, 1 y, 2 z; the low face if forward, else the high face), p = 0 on the
    opposite face, and the linear interpolant of the two on the four lateral
    faces, evaluated at the face centres (cell-centre coordinates along the
    axis)."""
    n_ax = (nx, ny, nz)[axis]
    frac = (np.arange(n_ax) + 0.5) / n_ax
    prof = 1.0 - frac if forward else frac            # value at cell index along the axis
    ii, jj, kk = np.arange(nx), np.arange(ny), np.arange(nz)
    # lateral faces: l/r over (j, k), u/d over (i, k), t/b over (i, j)
    grids = {
        "lr": np.meshgrid(jj, kk, indexing="xy"),     # arrays [k][j] -> flattened k*ny + j
        "ud": np.meshgrid(ii, kk, indexing="xy"),     # [k][i]
        "tb": np.meshgrid(ii, jj, indexing="xy"),     # [j][i]
    }
    coord = {  # which face-array coordinate is the axis' index, per face family
        "lr": {1: 0, 2: 1}, "ud": {0: 0, 2: 1}, "tb": {0: 0, 1: 1},
    }
    faces = [None] * 6
    fam_axis = {"lr": 0, "ud": 1, "tb": 2}
    fam_faces = {"lr": (0, 1), "ud": (3, 2), "tb": (5, 4)}     # (low face, high face)
    for fam, ax in fam_axis.items():
        lo, hi = fam_faces[fam]
        size = face_index_sizes(nx, ny, nz)[lo]
        if ax == axis:
            faces[lo] = np.full(size, 1.0 if forward else 0.0)
            faces[hi] = np.full(size, 0.0 if forward else 1.0)
        else:
            a, b = grids[fam]
            idx = (a, b)[coord[fam][axis]]
            v = prof[idx].ravel()
            faces[lo] = v.copy()
            faces[hi] = v.copy()
    return [np.ascontiguousarray(f, np.float64) for f in faces]


def source_sink_pair(nx, ny, nz, seed, volume):
    """b-only source term of a unit source and a unit sink in two distinct
    seeded cells: +1/volume and -1/volume (q/V)."""
    st = Stream(seed, 0)
    n = nx * ny * nz
    a = int(st.uniform() * n) % n
    b = a
    while b == a:
        b = int(st.uniform() * n) % n
    q = np.zeros(n)
    q[a] = 1.0 / volume
    q[b] = -1.0 / volume
    return q
