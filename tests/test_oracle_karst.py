"""Pin the oracle's 1-D conduit model and coupling loop against the reference.

tests/golden/karst/<case>.npz was written by the reference build running the
reference's own reservoir(), conduit() (permeabilidade_karst,
source_term1d, solve_linsys_1d + compute), fluxo1d() and fluxos() in
main.c's fixed-point loop (tests/golden/make_golden.py karst). The oracle
(oracle/hmfem_oracle.c driven by tests/oracle_py.coupled) must reproduce
them bit for bit: conduit pressures before the loop, after its first pass and
at the end, conduit fluxes, breccia velocities, the reservoir pressure and
fluxes (by digest), the per-pass changes and every solve's iteration count.
"""
import glob
import os

import numpy as np
import pytest

import oracle_py as O
from summary_sample import conduit_cells, field_digest, karst_sample_index

KARST = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "karst", "*.npz")))
SMALL = [p for p in KARST if "mixed" in p]
BIG = [p for p in KARST if "mixed" not in p]


def _names(paths):
    return [os.path.splitext(os.path.basename(p))[0] for p in paths]


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def test_fixtures_present():
    assert len(SMALL) >= 1 and len(BIG) >= 1


@pytest.mark.parametrize("path", SMALL, ids=_names(SMALL))
def test_coupled_loop_bitexact(path):
    z = np.load(path)
    r = O.coupled(z)
    assert r["passes"] == int(z["passes"])
    assert np.array_equal(r["pass_change"].view(np.uint64), _bits(z["pass_change"]))
    assert np.array_equal(r["iter_amg"], z["iter_amg"])
    for k in ("pc0", "pc1", "pc", "ucl", "ucr", "ucR_old", "ucL_old", "kc"):
        assert np.array_equal(_bits(r[k]), _bits(z[k])), k
    assert field_digest(r["p"]) == str(z["p_sha256"])
    assert field_digest(r["p1"]) == str(z["p1_sha256"])
    assert field_digest(r["v"].view(np.float64).reshape(-1)) == str(z["v_sha256"])
    idx = karst_sample_index(tuple(z["dims"]), z["start"], z["axis"], z["cells"])
    assert np.array_equal(_bits(r["p"][idx]), _bits(z["p_s"]))


@pytest.mark.parametrize("path", BIG, ids=_names(BIG))
def test_first_pass_bitexact(path):
    """The reference's own conduit runs (210x42x20): the loop's first pass."""
    z = np.load(path)
    r = O.coupled(z, max_passes=1)
    assert np.array_equal(r["pass_change"].view(np.uint64), _bits(z["pass_change"][:1]))
    assert np.array_equal(r["iter_amg"], z["iter_amg"][:2])
    for k in ("pc0", "pc1"):
        assert np.array_equal(_bits(r[k]), _bits(z[k])), k
    assert field_digest(r["p1"]) == str(z["p1_sha256"])


@pytest.mark.parametrize("path", SMALL, ids=_names(SMALL))
def test_direct_1d_solve_matches_tight_amg(path):
    """The oracle's exact tridiagonal solve (the GPU's comparator) agrees with
    the reference's AMG-PCG on the same 1-D systems driven to 1e-13, and is
    within the reference's own 1e-8 stop of its 1e-8 result."""
    z = np.load(path)
    d = O.coupled(z, max_passes=1, direct=True)
    t = O.coupled(z, max_passes=1, eps1d=1e-13)
    r = O.coupled(z, max_passes=1)
    for k in ("pc0", "pc1"):
        assert np.abs(d[k] - t[k]).max() <= 1e-12 * np.abs(t[k]).max(), k
        assert np.abs(d[k] - r[k]).max() <= 1e-7 * np.abs(r[k]).max(), k


def test_conduit_cells_and_overlap_order():
    """Conduit cell indices follow main.c:404-442 (start plus offset along
    the axis; x fastest)."""
    cells = conduit_cells((10, 8, 6), [1, 2, 3, 4, 0, 1], [1, 3], [3, 2])
    assert cells.tolist() == [1 + 10 * (2 + 8 * 3), 2 + 10 * (2 + 8 * 3), 3 + 10 * (2 + 8 * 3),
                              4 + 10 * (0 + 8 * 1), 4 + 10 * (0 + 8 * 2)]
