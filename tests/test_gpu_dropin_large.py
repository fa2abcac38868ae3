"""The host-buffer drop-ins on a grid whose AoS arrays span several pinned
staging chunks (96 x 96 x 72: perm is 37 MB, more than one 32 MB chunk):
permeabilidade_arestas and fluxos bit for bit against the oracle (fluxos
leaving every c as it was), solve_linsys to its stopping test."""
import numpy as np
import pytest

import oracle_py as O
from gpu_helpers import face_arrays, make_problem
from hmfem_amd import solver as S

pytestmark = pytest.mark.gpu


def test_dropins_across_staging_chunks():
    dims = (96, 96, 72)
    P = make_problem(dims, "lognormal", seed=5, conduits=4)
    n = P.n
    N = S.make_dim(P.nx, P.ny, P.nz)
    print("[dropin-large]", dims, flush=True)
    perm = np.zeros(n, S.ELE)
    indcol = np.zeros(n, S.IND)
    S.permeabilidade_arestas(P.Kx, P.Ky, P.Kz, perm, indcol, P.h, N, P.bc_type, P.WI, P.KI, P.visc)
    operm, oind, ob = O.assemble(P)
    for f in ("u", "d", "l", "r", "c", "t", "b"):
        assert np.array_equal(perm[f].view(np.uint64), operm[f].view(np.uint64)), f
        assert np.array_equal(indcol[f], oind[f]), f
    faces = face_arrays(P)
    b = np.zeros(n)
    S.source_term(faces, perm, P.h, N, P.bc_type, P.pc, P.pw, P.KI, P.WI, P.visc, b)
    assert np.array_equal(b.view(np.uint64), ob.view(np.uint64))
    x = np.zeros(n)
    S.solve_linsys(perm, indcol, b.copy(), x, N)
    rel = np.linalg.norm(b - O.prod_mat_vec(perm, indcol, x)) / np.linalg.norm(b)
    assert rel < 1e-6, rel
    S.set_boundary(P.bc_type)
    v = np.zeros(n, S.ELE)
    v["c"] = np.arange(n) * 0.25 - 3.0
    S.fluxos(x, perm, v, faces, P.h, N)
    ov = O.fluxos(P, x, operm)
    for f in ("u", "d", "l", "r", "t", "b"):
        assert np.array_equal(v[f].view(np.uint64), ov[f].view(np.uint64)), f
    assert np.array_equal(v["c"], np.arange(n) * 0.25 - 3.0)
