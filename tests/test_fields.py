"""Seeded permeability generators: determinism and statistics."""
import numpy as np

from hmfem_amd import fields


def test_lognormal_deterministic_and_standardised():
    a = fields.lognormal(24, 20, 16, np.log(1e-14), 2.0, 4.0, 7)
    b = fields.lognormal(24, 20, 16, np.log(1e-14), 2.0, 4.0, 7)
    c = fields.lognormal(24, 20, 16, np.log(1e-14), 2.0, 4.0, 8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    g = (np.log(a) - np.log(1e-14)) / 2.0
    assert abs(g.mean()) < 1e-12 and abs(g.std() - 1.0) < 1e-12


def test_lognormal_correlation_length():
    """Sample correlation at lag l along x is near exp(-1) for ell = l."""
    ell = 6.0
    K = fields.lognormal(96, 64, 64, 0.0, 1.0, ell, 3).reshape(64, 64, 96)
    g = np.log(K)
    lag = int(ell)
    c = np.mean(g[:, :, :-lag] * g[:, :, lag:]) / np.mean(g * g)
    assert 0.2 < c < 0.55


def test_conduits_face_connected_and_boosted():
    nx = ny = nz = 48
    K = fields.conduits(fields.homogeneous(nx, ny, nz, 1e-14), nx, ny, nz, 3, 11)
    m = K / 1e-14
    boosted = m > 1.5
    assert boosted.any()
    assert m[boosted].min() >= 1e4 * (1 - 1e-6) and m.max() <= 1e6 * (1 + 1e-6)
    # every boosted cell has a boosted face neighbour (paths are face-connected)
    B = boosted.reshape(nz, ny, nx)
    nb = np.zeros_like(B)
    nb[:, :, 1:] |= B[:, :, :-1]; nb[:, :, :-1] |= B[:, :, 1:]
    nb[:, 1:, :] |= B[:, :-1, :]; nb[:, :-1, :] |= B[:, 1:, :]
    nb[1:, :, :] |= B[:-1, :, :]; nb[:-1, :, :] |= B[1:, :, :]
    assert np.all(nb[B])


def test_instance_stream():
    s1 = fields.Stream(5, 3)
    s2 = fields.Stream(5, 3)
    a = [s1.uniform() for _ in range(5)]
    assert a == [s2.uniform() for _ in range(5)]
    assert all(0.0 <= x < 1.0 for x in a)
