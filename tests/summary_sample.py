"""Cells at which the full-size summary fixtures keep the pressure.

Every 8th cell along each axis (i, j, k multiples of 8) plus the three full
mid-planes i = nx/2, j = ny/2, k = nz/2; index s = i + nx*(j + ny*k),
ascending. Shared by the fixture writer and the tests that read it.
"""
import numpy as np


def sample_index(dims, stride=8):
    nx, ny, nz = dims
    i = np.arange(nx)[None, None, :]
    j = np.arange(ny)[None, :, None]
    k = np.arange(nz)[:, None, None]
    keep = ((i % stride == 0) & (j % stride == 0) & (k % stride == 0)) | (i == nx // 2) | (j == ny // 2) | (k == nz // 2)
    return np.flatnonzero(keep.ravel()).astype(np.int64)


def field_digest(K):
    """SHA-256 of the FP64 bytes of a field: ties a fixture to the field it was made from."""
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(K, "<f8").tobytes()).hexdigest()
