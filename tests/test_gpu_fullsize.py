"""Full-size pressures against the reference's own solutions.

tests/golden/full/<case>.npz holds what the reference build (the reference's
own assembly, count-CSR, AMG-PCG and flux sources, compiled from source)
computed on the two single-grid benchmark workloads:

  c2_128_lognormal  128^3, ln K ~ N(ln 1e-14, 2^2), correlation 8 cells
  c3_256_karst      256^3, K = 1e-14 plus 20 seeded conduits x10^U(4,6)

both with p = 1 / p = 0 on the x faces and no flow elsewhere: the pressure
at the reference's own stop (compute.cpp:41, pcg.h:57, 1e-8), after a
warm-started second pass, and driven to 1e-12, at the cells of
tests/summary_sample.py, plus whole-field norms, sums and iteration counts.

Bars:
  * driven to a true relative residual of 1e-12, the GPU pressure is within
    1e-8 relative L2 of the reference's 1e-12 pressure;
  * at the reference's own stopping rule and tolerance (pcg.h:57, 1e-8) the
    GPU pressure is at least as close to the reference's 1e-12 pressure as
    the reference's own 1e-8 iterate is; at the benchmark's rule (true
    relative residual 1e-8) within twice that distance (a residual bound
    does not bound the error, and the GPU's preconditioner reaches the
    residual in fewer iterations);
  * a tolerance down to 1e-10 puts the GPU pressure within 1e-8 relative L2
    of the reference's 1e-12 pressure (the reference's own 1e-8 iterate is
    5.6e-8 away on the 128^3 case and 6.5e-7 on the 256^3 conduit case);
  * iteration ceilings (a regression that slows the preconditioner fails).
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle_py as O
from gpu_helpers import GOLD, Dev, torch_cuda
from hmfem_amd import fields as F
from hmfem_amd import solver as S
from summary_sample import field_digest, sample_index

pytestmark = pytest.mark.gpu

FULL = os.path.join(GOLD, "full")
CASES = ["c2_128_lognormal", "c3_256_karst"]
# CG iterations of the default FP32 hierarchy at rtol 1e-8 (true residual):
# 17 on the 128^3 case and 69 on the 256^3 conduit case when these ceilings
# were set; the ceilings leave room for summation-order changes only.
ITER_CEILING_1E8 = {"c2_128_lognormal": 19, "c3_256_karst": 78}


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


def load_case(name):
    z = np.load(os.path.join(FULL, name + ".npz"))
    dims = tuple(int(v) for v in z["dims"])
    spec = json.loads(str(z["K_spec"]))
    nx, ny, nz = dims
    if spec[0] == "lognormal":
        K = F.lognormal(nx, ny, nz, np.log(1e-14), spec[1], spec[2], spec[3])
    else:
        K = F.conduits(F.homogeneous(nx, ny, nz, 1e-14), nx, ny, nz, spec[1], spec[2])
    assert field_digest(K) == str(z["K_sha256"]), "the field generator no longer makes the fixture's K"
    P = O.Problem(dims, K, tuple(z["bc_type"]), tuple(z["bc_val"]), float(z["visc"]))
    idx = sample_index(dims)
    assert idx.size == int(z["sample_idx_count"])
    return z, P, idx


def sampled_rel(p, idx, ref_s):
    return float(np.linalg.norm(p[idx] - ref_s) / np.linalg.norm(ref_s))


def solve(ctx, torch, P, **kw):
    D = Dev(P, torch)
    p = torch.zeros(P.n, dtype=torch.float64, device="cuda")
    v = torch.zeros(P.n * 7, dtype=torch.float64, device="cuda")
    st = ctx.reservoir(D.g, D.K, D.K, D.K, D.bval, p, v, S.default_options(**kw))
    x = p.cpu().numpy()
    del p, v, D
    torch.cuda.empty_cache()
    return st, x


@pytest.mark.parametrize("name", CASES)
def test_tight_pressure_matches_reference(name, ctx, torch):
    z, P, idx = load_case(name)
    _say("full-size tight", name)
    st, x = solve(ctx, torch, P, rtol=1e-12)
    err = sampled_rel(x, idx, z["p12_s"])
    _say(name, "rtol 1e-12:", st.as_dict(), "sampled rel L2 vs reference p12 %.3e" % err,
         "norm %.17g vs %.17g, sum %.17g vs %.17g, div_max %.3e vs %.3e"
         % (np.linalg.norm(x), z["p12_norm2"], x.sum(), z["p12_sum"], st.div_max, z["div_max"]))
    assert st.status == S.OK
    assert err < 1e-8
    assert abs(np.linalg.norm(x) - z["p12_norm2"]) <= 1e-8 * z["p12_norm2"]
    assert abs(x.sum() - z["p12_sum"]) <= 1e-8 * abs(z["p12_sum"])


@pytest.mark.parametrize("name", CASES)
def test_benchmark_tolerance_vs_reference(name, ctx, torch):
    """Under the reference's stopping rule at 1e-8 the GPU pressure is no
    further from the reference's converged pressure than the reference's own
    iterate at its own 1e-8 stop; under the benchmark's true-residual rule at
    1e-8, within twice that; a tolerance down to 1e-10 meets the 1e-8 bar."""
    z, P, idx = load_case(name)
    ref_own = float(z["rel_l2_p8_p12"])
    ref_own_s = sampled_rel(z["p8_s"], np.arange(idx.size), z["p12_s"])
    st, x = solve(ctx, torch, P, rtol=1e-8)
    err = sampled_rel(x, idx, z["p12_s"])
    st_p, xp = solve(ctx, torch, P, rtol=1e-8, stop_rule=S.STOP_PRECONDITIONED)
    err_p = sampled_rel(xp, idx, z["p12_s"])
    bar = None
    for rt in (1e-9, 1e-10):
        st_t, x_t = solve(ctx, torch, P, rtol=rt)
        e_t = sampled_rel(x_t, idx, z["p12_s"])
        _say(name, "rtol %g: %d its, err %.3e" % (rt, st_t.iterations, e_t))
        if st_t.status == S.OK and e_t < 1e-8:
            bar = rt
            break
    _say(name, "rtol 1e-8: %d its, err %.3e; preconditioned rule: %d its, err %.3e; 1e-8 bar met at rtol %s; "
         "reference at its 1e-8 stop: %d loops, err %.3e (sampled %.3e)"
         % (st.iterations, err, st_p.iterations, err_p, bar, int(z["iter8"]), ref_own, ref_own_s))
    assert st.status == S.OK and st_p.status == S.OK
    assert err <= 2 * ref_own_s
    assert err_p <= ref_own_s
    assert bar is not None
    assert st.iterations <= ITER_CEILING_1E8[name]
