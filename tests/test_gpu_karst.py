"""The coupled reservoir / 1-D conduit model on the GPU (hmfem_coupled_device)
against the reference and the oracle.

tests/golden/karst/<case>.npz: the reference build running main.c's loop
(reservoir() and conduit() until the reservoir pressure changes by at most
1e-8 between passes, then fluxo1d() and fluxos()) on the reference's own
setup (K = 1e-13, mu = 1e-3, 60 MPa lateral faces, a well column) with the
reference's conduit data files and a mixed set of every end type and axis.

The GPU solves each conduit's tridiagonal system exactly where the
reference's compute() stops its AMG-PCG at 1e-8, and its reservoir
preconditioner is its own, so the bars are tolerances:
  * converged loop: conduit and reservoir pressures within 1e-7 relative of
    the reference's, the same pass count within a few passes;
  * the oracle with the same exact 1-D solve and both solves driven tight,
    for a fixed number of passes: pressures within 1e-10, conduit fluxes
    within 1e-8 relative.
"""
import glob
import os
import sys

import numpy as np
import pytest

import oracle_py as O
from gpu_helpers import GOLD, torch_cuda
from hmfem_amd import solver as S
from summary_sample import karst_sample_index

pytestmark = pytest.mark.gpu

KARST = sorted(glob.glob(os.path.join(GOLD, "karst", "*.npz")))


def _names(paths):
    return [os.path.splitext(os.path.basename(p))[0] for p in paths]


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


def conduit_set(z):
    bc = np.asarray(z["bc"]).reshape(-1, 2)
    return S.ConduitSet(z["cells"], z["start"], z["axis"], z["Kc"], bc[:, 0], bc[:, 1], z["cont"], z["ccont"],
                        z["karst_index"], z["area"], z["nbrec"], z["brec"], z["beta"], z["C"], float(z["rho"]))


def flat_conduits(a, cells):
    """[row][conduit] matrix (the reference's layout) -> conduits one after the other."""
    return np.concatenate([a[:m, j] for j, m in enumerate(cells)])


def run_gpu(ctx, torch, z, max_passes=1000, tol=1e-8, rtol=1e-8):
    P, WI = O.karst_reservoir(z)
    cs = conduit_set(z)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to("cuda")
    nx, ny, nz = P.nx, P.ny, P.nz
    sizes = (ny * nz, ny * nz, nx * nz, nx * nz, nx * ny, nx * ny)
    bval = [dev(np.full(sizes[f], P.bc_val[f])) for f in range(6)]
    K = dev(P.Kx)
    g = S.Context.grid(nx, ny, nz, P.h, P.visc, P.bc_type)
    p = torch.zeros(P.n, dtype=torch.float64, device="cuda")
    nc = cs.total_cells
    pc, ucl, ucr, pc0 = (torch.zeros(nc, dtype=torch.float64, device="cuda") for _ in range(4))
    v = torch.zeros(P.n * 7, dtype=torch.float64, device="cuda")
    opt = S.default_options(rtol=rtol, stop_rule=S.STOP_PRECONDITIONED, reuse_setup=1)
    st, changes = ctx.coupled(g, K, K, K, bval, cs, p, pc, WI=dev(WI), pw=dev(P.pw), max_passes=max_passes, tol=tol,
                              ucl=ucl, ucr=ucr, v=v, pc_first=pc0, options=opt)
    out = dict(p=p.cpu().numpy(), pc=pc.cpu().numpy(), ucl=ucl.cpu().numpy(), ucr=ucr.cpu().numpy(),
               pc0=pc0.cpu().numpy(), v=v.cpu().numpy().view(S.ELE))
    return st, changes, out


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("path", KARST, ids=_names(KARST))
def test_coupled_loop_vs_reference(path, ctx, torch):
    z = np.load(path)
    _say("coupled loop", os.path.basename(path))
    st, changes, g = run_gpu(ctx, torch, z)
    cells = [int(m) for m in z["cells"]]
    idx = karst_sample_index(tuple(z["dims"]), z["start"], z["axis"], z["cells"])
    e_p = rel(g["p"][idx], z["p_s"])
    e_pc = rel(g["pc"], flat_conduits(z["pc"], cells))
    e_pc0 = rel(g["pc0"], flat_conduits(z["pc0"], cells))
    e_ucl = rel(g["ucl"], flat_conduits(z["ucl"], cells))
    e_ucr = rel(g["ucr"], flat_conduits(z["ucr"], cells))
    _say(os.path.basename(path), st.as_dict(), "reference passes %d; rel diff: p %.3e pc %.3e pc0 %.3e ucl %.3e ucr %.3e"
         % (int(z["passes"]), e_p, e_pc, e_pc0, e_ucl, e_ucr))
    assert st.status == S.OK
    assert abs(st.passes - int(z["passes"])) <= max(3, int(z["passes"]) // 10)
    assert e_p < 1e-7 and e_pc < 1e-7 and e_pc0 < 1e-7
    assert abs(np.linalg.norm(g["p"]) - float(z["p_norm2"])) <= 1e-7 * float(z["p_norm2"])
    assert e_ucl < 1e-4 and e_ucr < 1e-4


@pytest.mark.parametrize("path", [p for p in KARST if "mixed" in p], ids=["mixed"])
def test_fixed_passes_vs_oracle_tight(path, ctx, torch):
    """Three passes with both solves driven tight: the GPU against the
    oracle with the same exact 1-D solve (pinned against the reference by
    tests/test_oracle_karst.py)."""
    z = np.load(path)
    n_pass = 3
    o = O.coupled(z, max_passes=n_pass, eps3d=1e-13, direct=True, tol=0.0)
    st, changes, g = run_gpu(ctx, torch, z, max_passes=n_pass, tol=0.0, rtol=1e-13)
    cells = [int(m) for m in z["cells"]]
    e = {k: rel(g[k], flat_conduits(o[k], cells)) for k in ("pc", "pc0", "ucl", "ucr")}
    e["p"] = rel(g["p"], o["p"])
    e["change"] = float(np.max(np.abs(changes - o["pass_change"]) / o["pass_change"]))
    _say("fixed passes vs oracle", st.as_dict(), e)
    assert st.passes == n_pass and o["passes"] == n_pass
    assert e["p"] < 1e-10 and e["pc"] < 1e-10 and e["pc0"] < 1e-10
    assert e["ucl"] < 1e-8 and e["ucr"] < 1e-8
    assert e["change"] < 1e-6
