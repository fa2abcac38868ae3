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


def conduit_cells(dims, start, axis, cells):
    """Reservoir cell index of every conduit cell, conduit by conduit
    (main.c:404-442: conduit j runs from (sxi, syi, szi) along axis 1 x,
    2 y, 3 z)."""
    nx, ny, nz = dims
    out = []
    for j, m in enumerate(cells):
        x, y, z = start[3 * j:3 * j + 3]
        o = np.arange(m)
        xs, ys, zs = (x + o, np.full(m, y), np.full(m, z)) if axis[j] == 1 else \
            (np.full(m, x), y + o, np.full(m, z)) if axis[j] == 2 else (np.full(m, x), np.full(m, y), z + o)
        out.append(xs + nx * (ys + ny * zs))
    return np.concatenate(out).astype(np.int64) if out else np.zeros(0, np.int64)


def karst_sample_index(dims, start, axis, cells):
    """The three mid-planes plus every conduit cell, ascending."""
    nx, ny, nz = dims
    i = np.arange(nx)[None, None, :]
    j = np.arange(ny)[None, :, None]
    k = np.arange(nz)[:, None, None]
    keep = ((i == nx // 2) | (j == ny // 2) | (k == nz // 2)).ravel()
    keep[conduit_cells(dims, start, axis, cells)] = True
    return np.flatnonzero(keep).astype(np.int64)
