"""hmfem_flux_device_noalloc against hmfem_flux_device, through the C ABI.

Two calls in a row on one context, the second on a smaller grid (the fixed
scratch then still holds the first call's partials beyond the second call's
range): fluxes and divergence min/max must equal the allocating entry
point's bit for bit. Also: a grid beyond the fixed scratch is rejected
before any device work.
"""
import sys

import numpy as np
import pytest

from gpu_helpers import Dev, make_problem, torch_cuda
from hmfem_amd import solver as S

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_noalloc_flux_matches_allocating_entry_point():
    torch = torch_cuda()
    ctx = S.Context()
    try:
        for dims, kind, seed in (((40, 36, 30), "lognormal", 5), ((17, 13, 11), "karst", 6)):
            _say("no-alloc flux", dims, kind)
            P = make_problem(dims, kind, seed=seed, corr=4.0, conduits=3)
            D = Dev(P, torch)
            p = torch.zeros(P.n, dtype=torch.float64, device="cuda")
            ctx.reservoir(D.g, D.K, D.K, D.K, D.bval, p, None, S.default_options(rtol=1e-10))
            v_ref = torch.full((P.n * 7,), 7.0, dtype=torch.float64, device="cuda")
            v_new = torch.full((P.n * 7,), 7.0, dtype=torch.float64, device="cuda")
            mn_ref, mx_ref = ctx.flux(D.g, D.K, D.K, D.K, D.bval, p, v_ref)
            mn_new, mx_new = ctx.flux_noalloc(D.g, D.K, D.K, D.K, D.bval, p, v_new)
            _say("  div min/max", mn_ref, mx_ref, mn_new, mx_new)
            assert torch.equal(v_ref.view(torch.int64), v_new.view(torch.int64))
            assert np.float64(mn_ref).view(np.uint64) == np.float64(mn_new).view(np.uint64)
            assert np.float64(mx_ref).view(np.uint64) == np.float64(mx_new).view(np.uint64)
            assert mx_new > 0
    finally:
        ctx.close()


def test_noalloc_flux_rejects_grids_beyond_its_scratch():
    torch = torch_cuda()
    ctx = S.Context()
    try:
        tiny = torch.zeros(8, dtype=torch.float64, device="cuda")
        g = S.Context.grid(1024, 1024, 129)   # 2^27 + 2^20 cells: one plane too many
        with pytest.raises(RuntimeError):
            ctx.flux_noalloc(g, tiny, tiny, tiny, [tiny] * 6, tiny, None)
    finally:
        ctx.close()
