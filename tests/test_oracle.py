"""Pin the CPU restatement (oracle/) before trusting it as the checker.

Every fixture in tests/golden/ was written by the reference's own sources
(tests/golden/make_golden.py over oracle/_ref). The restatement must match
them bit for bit; closed forms and brute force pin what fixtures cannot.
"""
import glob
import os

import numpy as np
import pytest

import oracle_py as O

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
FIELDS = ("u", "d", "l", "r", "t", "b")


def _names(paths):
    return [os.path.splitext(os.path.basename(p))[0] for p in paths]


@pytest.fixture(params=GOLD, ids=_names(GOLD))
def fixture(request):
    return np.load(request.param)


def test_fixtures_present():
    assert len(GOLD) >= 6


def test_assembly_and_rhs_bitexact(fixture):
    P = O.problem_from_fixture(fixture)
    perm, indcol, b = O.assemble(P)
    assert np.array_equal(b.view(np.uint64), fixture["b"].view(np.uint64))
    if "cnt" in fixture.files:
        cnt, col, ele = O.build_csr(perm, indcol)
        assert np.array_equal(cnt, fixture["cnt"])
        assert np.array_equal(col, fixture["col"])
        assert np.array_equal(ele.view(np.uint64), fixture["ele"].view(np.uint64))


def test_compute_reference_stop_bitexact(fixture):
    """compute() at the reference's 1e-8 stop: same iterate, same _loop."""
    P = O.problem_from_fixture(fixture)
    perm, indcol, b = O.assemble(P)
    cnt, col, ele = O.build_csr(perm, indcol)
    x, it = O.compute(cnt, col, ele, b, eps=1e-8)
    assert it == int(fixture["iter8"])
    assert np.array_equal(x.view(np.uint64), fixture["p8"].view(np.uint64))


def test_compute_warm_start_bitexact(fixture):
    """A second compute() started from the first pass's pressure, as
    main.c:460 hands p back: the stop test is relative to the warm start's
    own preconditioned residual (pcg.h:42-57), so the pass keeps iterating."""
    P = O.problem_from_fixture(fixture)
    perm, indcol, b = O.assemble(P)
    cnt, col, ele = O.build_csr(perm, indcol)
    x, it = O.compute(cnt, col, ele, b, x0=fixture["p8"], eps=1e-8)
    assert it == int(fixture["iter8w"])
    assert np.array_equal(x.view(np.uint64), fixture["p8w"].view(np.uint64))


def test_tight_solve_levels_and_flux_bitexact(fixture):
    P = O.problem_from_fixture(fixture)
    perm, indcol, b = O.assemble(P)
    cnt, col, ele = O.build_csr(perm, indcol)
    amg = O.AMG(cnt, col, ele)
    rows = [r for r, _ in amg.levels()]
    nnzs = [z for _, z in amg.levels()]
    assert rows == list(fixture["levels_rows"])
    assert nnzs == list(fixture["levels_nnz"])
    x, it = amg.pcg(b, eps=1e-12)
    assert it == int(fixture["iter12"])
    assert np.array_equal(x.view(np.uint64), fixture["p12"].view(np.uint64))
    v = O.fluxos(P, x, perm)
    if "v" in fixture.files:
        ref = fixture["v"]
        for k, f in enumerate(("u", "d", "l", "r", "c", "t", "b")):
            if f == "c":
                continue
            assert np.array_equal(v[f].view(np.uint64), ref[:, k].view(np.uint64)), f
    dmin, dmax = O.divvel_pontual(P, v)
    assert dmax == float(fixture["div_max"]) and dmin == float(fixture["div_min"])


def _c1():
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_16_homogeneous.npz"))
    return z, O.problem_from_fixture(z)


def test_c1_closed_form_pressure_and_flux():
    """16^3, K=1e-14, lambda=1, p=1 at x=0, p=0 at x=1: p_i = 1 - (i+1/2)/16
    exactly for this scheme, flux K*lambda through every x face, div = 0."""
    z, P = _c1()
    perm, indcol, b = O.assemble(P)
    cnt, col, ele = O.build_csr(perm, indcol)
    x, _ = O.AMG(cnt, col, ele).pcg(b, eps=1e-12)
    i = np.arange(P.n) % 16
    exact = 1.0 - (i + 0.5) / 16.0
    assert np.linalg.norm(x - exact) / np.linalg.norm(exact) < 1e-10
    v = O.fluxos(P, x, perm)
    Kl = 1e-14
    assert np.allclose(v["r"], Kl, rtol=1e-9, atol=0)
    assert np.allclose(v["l"], -Kl, rtol=1e-9, atol=0)
    assert np.max(np.abs(v["u"])) < 1e-9 * Kl and np.max(np.abs(v["t"])) < 1e-9 * Kl
    dmin, dmax = O.divvel_pontual(P, v)
    assert dmax < 1e-9 * Kl * 16


def test_transmissibility_closed_forms():
    """Uniform K: interior T = K/(h^2 mu), Dirichlet half-cell T = 2K/(h^2 mu),
    Neumann T = 0, centre = sum of faces."""
    nx, ny, nz = 4, 3, 5
    K = np.full(nx * ny * nz, 3e-13)
    P = O.Problem((nx, ny, nz), K, (0, 1, 1, 0, 1, 1), (1, 0, 0, 0, 0, 0), visc=2.0)
    perm, indcol, _ = O.assemble(P)
    hx, hy, hz = P.h
    s = 1 + nx * 1 + nx * ny * 2  # interior cell (1,1,2)
    assert np.isclose(perm["r"][s], 3e-13 / (hx * hx * 2.0), rtol=1e-15)
    assert np.isclose(perm["u"][s], 3e-13 / (hy * hy * 2.0), rtol=1e-15)
    assert np.isclose(perm["t"][s], 3e-13 / (hz * hz * 2.0), rtol=1e-15)
    s0 = 0
    assert np.isclose(perm["l"][s0], 2 * 3e-13 / (hx * hx * 2.0), rtol=1e-15) and indcol["l"][s0] == -1
    assert np.isclose(perm["d"][s0], 2 * 3e-13 / (hy * hy * 2.0), rtol=1e-15) and indcol["d"][s0] == -1
    assert perm["b"][s0] == 0.0 and indcol["b"][s0] == -1
    faces = sum(perm[f] for f in FIELDS)
    assert np.allclose(perm["c"], faces, rtol=1e-15)


def test_matrix_symmetric_m_matrix(fixture):
    if "cnt" not in fixture.files:
        pytest.skip("fixture keeps no matrix")
    n = int(np.prod(fixture["dims"]))
    cnt, col, ele = fixture["cnt"], fixture["col"], fixture["ele"]
    rows = np.repeat(np.arange(n), cnt)
    A = np.zeros((n, n)) if n <= 4096 else None
    if A is None:
        pytest.skip("too large for a dense check")
    A[rows, col] = ele
    assert np.array_equal(A, A.T)                      # bitwise symmetric (hm is symmetric)
    off = A - np.diag(np.diag(A))
    assert (off <= 0).all()                            # M-matrix sign pattern
    assert (np.diag(A) + off.sum(1) >= -1e-12 * np.abs(np.diag(A))).all()  # weak diagonal dominance


def test_galerkin_brute_force():
    """A_c = P^T A P (dense), P rows sum to 1, A_c symmetric, on every level."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "lognormal_8.npz"))
    amg = O.AMG(z["cnt"], z["col"], z["ele"])
    lv = amg.levels()
    for l in range(len(lv) - 1):
        n, nc = lv[l][0], lv[l + 1][0]
        cnt, col, ele = amg.level_matrix(l)
        A = np.zeros((n, n)); A[np.repeat(np.arange(n), cnt), col] = ele
        pc, pcol, pele = amg.level_interp(l)
        Pm = np.zeros((n, nc)); Pm[np.repeat(np.arange(n), pc), pcol] = pele
        assert np.allclose(Pm.sum(1), 1.0, rtol=1e-15)
        cc, ccol, cele = amg.level_matrix(l + 1)
        Ac = np.zeros((nc, nc)); Ac[np.repeat(np.arange(nc), cc), ccol] = cele
        ref = Pm.T @ A @ Pm
        scale = np.abs(ref).max()
        assert np.abs(Ac - ref).max() <= 1e-13 * scale
        assert np.abs(Ac - Ac.T).max() <= 1e-13 * scale


def test_pcg_matches_dense_solve():
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "sources_10x8x6.npz"))
    n = int(np.prod(z["dims"]))
    A = np.zeros((n, n)); A[np.repeat(np.arange(n), z["cnt"]), z["col"]] = z["ele"]
    exact = np.linalg.solve(A, z["b"])
    assert np.linalg.norm(z["p12"] - exact) / np.linalg.norm(exact) < 1e-10


def test_planted_mistake_is_caught():
    """A one-ulp change in one transmissibility must change the bits of the
    solution: the bit-exact comparisons above are sensitive."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "lognormal_8.npz"))
    ele = z["ele"].copy()
    k = int(np.argmax(z["col"] != np.repeat(np.arange(512), z["cnt"])))
    ele[k] = np.nextafter(ele[k], 0.0)
    x, _ = O.AMG(z["cnt"], z["col"], ele).pcg(z["b"], eps=1e-12)
    assert not np.array_equal(x.view(np.uint64), z["p12"].view(np.uint64))
