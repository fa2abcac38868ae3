"""The GPU hierarchy and the preconditioner's quality.

* Every level the GPU builds is the Galerkin product of the level above it:
  A_l = P^T A_{l-1} P with P the piecewise-constant interpolation of the box
  aggregates (fine cell (i, j, k) -> coarse cell (i/fx, j/fy, k/fz)), the
  reference's triple product (amg.h:567-582) for the build's aggregates,
  checked against dense products on an odd grid; the finest level is the
  oracle's assembly bit for bit.
* Iteration ceilings at the default options on the small benchmark
  problems (the full-size ones are in tests/test_gpu_fullsize.py), and the
  conduit correction's gain at the shipped defaults on the 256^3 conduit
  case.
"""
import sys

import numpy as np
import pytest

import oracle_py as O
from gpu_helpers import Dev, make_problem, torch_cuda
from hmfem_amd import fields as F
from hmfem_amd import solver as S

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    return torch_cuda()


@pytest.fixture(scope="module")
def ctx(torch):
    c = S.Context()
    yield c
    c.close()


def dense(dims, e, wx, wy, wz):
    """The flux-form operator (A u)_s = e_s u_s + sum_f w_f (u_s - u_nb) as a dense matrix."""
    nx, ny, nz = dims
    n = nx * ny * nz
    A = np.diag(e.copy())
    for w, step, ext, stride in ((wx, 1, nx, 1), (wy, nx, ny, 1), (wz, nx * ny, nz, 1)):
        for s in range(n):
            c = (s // step) % ext
            if c < ext - 1 and w[s] != 0.0:
                t = s + step
                A[s, s] += w[s]
                A[t, t] += w[s]
                A[s, t] -= w[s]
                A[t, s] -= w[s]
    return A


def box_interpolation(fine, fac, coarse):
    nx, ny, nz = fine
    cx, cy, cz = coarse
    P = np.zeros((nx * ny * nz, cx * cy * cz))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                P[i + nx * (j + ny * k), i // fac[0] + cx * (j // fac[1] + cy * (k // fac[2]))] = 1.0
    return P


@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_FP64], ids=["fp32", "fp64"])
def test_levels_are_galerkin_products(prec, ctx, torch):
    nx, ny, nz = 13, 11, 9
    K = F.conduits(F.lognormal(nx, ny, nz, np.log(1e-14), 2.0, 3.0, 5), nx, ny, nz, 3, 8)
    P = O.Problem((nx, ny, nz), K, (0, 0, 1, 1, 1, 1), (1.0, 0.0, 0.0, 0.0, 0.0, 0.0))
    D = Dev(P, torch)
    p = torch.zeros(P.n, dtype=torch.float64, device="cuda")
    o = S.default_options(rtol=1e-10, precision=prec, coarse_max=8, tail_max=0)
    st = ctx.reservoir(D.g, D.K, D.K, D.K, D.bval, p, None, o)
    assert st.status == S.OK
    levels = ctx.hierarchy()
    _say("hierarchy", [(d, f) for d, f, *_ in levels])
    assert len(levels) == st.levels and len(levels) >= 3
    # the finest level: the oracle's assembly in flux form, bit for bit
    perm, indcol, _ = O.assemble(P)
    d0, f0, e, wx, wy, wz = levels[0]
    assert d0 == (nx, ny, nz)
    bnd = lambda f, idx: np.where(indcol[idx] == -1, perm[f], 0.0)
    e_ref = bnd("l", "l") + bnd("r", "r") + bnd("u", "u") + bnd("d", "d") + bnd("t", "t") + bnd("b", "b")
    assert np.array_equal(e.view(np.uint64), e_ref.view(np.uint64))
    for w, f in ((wx, "r"), (wy, "u"), (wz, "t")):
        assert np.array_equal(w.view(np.uint64), np.where(indcol[f] != -1, perm[f], 0.0).view(np.uint64))
    # every coarser level: P^T A P of the one above
    for l in range(1, len(levels)):
        df, ff, *cf = levels[l - 1]
        dc, _, *cc = levels[l]
        assert dc == tuple(-(-a // b) for a, b in zip(df, ff)), (l, df, ff, dc)
        Af = dense(df, *cf)
        Ac = dense(dc, *cc)
        Pm = box_interpolation(df, ff, dc)
        G = Pm.T @ Af @ Pm
        err = np.abs(Ac - G).max() / np.abs(G).max()
        _say("level", l, dc, "max |A_c - P^T A P| / max |A| = %.3e" % err)
        assert err < 1e-14, l
        assert np.allclose(Ac, Ac.T, rtol=0, atol=1e-15 * np.abs(Ac).max())
    # the coarsest level's dense inverse (block Gauss-Jordan on the matrix cores)
    inv = ctx.coarse_inverse()
    Ac = dense(levels[-1][0], *levels[-1][2:])
    assert inv.shape == Ac.shape
    ref = np.linalg.inv(Ac)
    e_inv = np.abs(inv - ref).max() / np.abs(ref).max()
    e_id = np.abs(inv @ Ac - np.eye(Ac.shape[0])).max()
    _say("coarsest %s: |inv - numpy inv| / |inv| = %.3e, |inv A - I| = %.3e" % (Ac.shape, e_inv, e_id))
    assert e_inv < 1e-12 and e_id < 1e-10


@pytest.mark.parametrize("dims", [(4, 4, 5), (3, 3, 3), (5, 4, 4), (2, 1, 1)], ids=["80", "27", "80b", "2"])
def test_coarse_inverse_sizes(dims, ctx, torch):
    """The dense coarsest inverse on levels of 80 (five tiles of 16), 27 and
    2 cells (padded to whole tiles): a grid that is its own coarsest level."""
    P = make_problem(dims, "lognormal", seed=3, sigma=1.5, corr=2.0)
    D = Dev(P, torch)
    p = torch.zeros(P.n, dtype=torch.float64, device="cuda")
    st = ctx.reservoir(D.g, D.K, D.K, D.K, D.bval, p, None, S.default_options(rtol=1e-12))
    assert st.status == S.OK and st.levels == 1
    lv = ctx.hierarchy()
    Ac = dense(lv[0][0], *lv[0][2:])
    inv = ctx.coarse_inverse()
    e_id = np.abs(inv @ Ac - np.eye(Ac.shape[0])).max()
    _say("coarsest", dims, "|inv A - I| = %.3e" % e_id)
    assert e_id < 1e-10


# CG iterations at the default options when these ceilings were set, with
# room for summation-order changes only: 16^3 homogeneous at 1e-10: 10;
# 32^3 lognormal sigma 2 at 1e-8: 10; 64^3 with 5 conduits at 1e-8: 31.
CEILINGS = [("c1_16_homogeneous", (16, 16, 16), "homogeneous", 1e-10, 12),
            ("lognormal_32", (32, 32, 32), "lognormal", 1e-8, 12),
            ("karst_64", (64, 64, 64), "karst", 1e-8, 36)]


@pytest.mark.parametrize("name,dims,kind,rtol,ceiling", CEILINGS, ids=[c[0] for c in CEILINGS])
def test_iteration_ceilings(name, dims, kind, rtol, ceiling, ctx, torch):
    P = make_problem(dims, kind, seed=20261202, sigma=2.0, corr=8.0 if dims[0] > 16 else 4.0, conduits=5)
    D = Dev(P, torch)
    p = torch.zeros(P.n, dtype=torch.float64, device="cuda")
    st = ctx.reservoir(D.g, D.K, D.K, D.K, D.bval, p, None, S.default_options(rtol=rtol))
    _say(name, "iterations", st.iterations, "ceiling", ceiling, "clusters", st.deflation_clusters)
    assert st.status == S.OK
    assert st.iterations <= ceiling


def test_conduit_correction_gain_at_defaults(ctx, torch):
    """The 256^3 conduit case at the shipped defaults: the correction on
    (the default) needs far fewer CG iterations than off, and the
    correction's solve is faster in wall time."""
    m = 256
    K = F.conduits(F.homogeneous(m, m, m, 1e-14), m, m, m, 20, 20261202)
    P = O.Problem((m, m, m), K, (0, 0, 1, 1, 1, 1), (1.0, 0.0, 0.0, 0.0, 0.0, 0.0))
    D = Dev(P, torch)
    p = torch.zeros(P.n, dtype=torch.float64, device="cuda")
    on = ctx.reservoir(D.g, D.K, D.K, D.K, D.bval, p, None, S.default_options())
    p.zero_()
    off = ctx.reservoir(D.g, D.K, D.K, D.K, D.bval, p, None, S.default_options(deflate=0))
    _say("256^3 conduits: correction on %d its %.1f ms (%d clusters), off %d its %.1f ms"
         % (on.iterations, on.t_setup_ms + on.t_solve_ms, on.deflation_clusters, off.iterations,
            off.t_setup_ms + off.t_solve_ms))
    assert on.status == S.OK and off.status == S.OK
    assert on.deflation_clusters >= 1 and off.deflation_clusters == 0
    assert on.iterations < 0.7 * off.iterations
    assert on.t_setup_ms + on.t_solve_ms < off.t_setup_ms + off.t_solve_ms
