"""Shared helpers for the GPU tests: problems, device copies, residuals."""
import os

import numpy as np

import oracle_py as O
from hmfem_amd import fields
from hmfem_amd import solver as S

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("these tests need a HIP device")
    return torch


def face_arrays(P):
    """The six Bound arrays of a problem, sized by face (l, r, u, d, t, b)."""
    nx, ny, nz = P.nx, P.ny, P.nz
    sizes = (ny * nz, ny * nz, nx * nz, nx * nz, nx * ny, nx * ny)
    return [np.ascontiguousarray(P.B[f][:sizes[f]]) for f in range(6)]


def make_problem(dims, kind, seed=1, sigma=2.0, corr=8.0, conduits=20, bc_type=(0, 0, 1, 1, 1, 1),
                 bc_val=(1.0, 0.0, 0.0, 0.0, 0.0, 0.0)):
    nx, ny, nz = dims
    if kind == "homogeneous":
        K = fields.homogeneous(nx, ny, nz, 1e-14)
    elif kind == "lognormal":
        K = fields.lognormal(nx, ny, nz, np.log(1e-14), sigma, corr, seed)
    elif kind == "karst":
        K = fields.conduits(fields.homogeneous(nx, ny, nz, 1e-14), nx, ny, nz, conduits, seed)
    else:
        raise ValueError(kind)
    return O.Problem(dims, K, bc_type, bc_val)


class Dev:
    """Device copies of a problem's inputs (torch tensors on cuda:0)."""

    def __init__(self, P, torch):
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
        self.P = P
        self.K = t(P.Kx)
        self.bval = [t(a) for a in face_arrays(P)]
        self.has_src = any(np.any(a != 0) for a in (P.KI, P.WI))
        self.KI = t(P.KI) if self.has_src else None
        self.WI = t(P.WI) if self.has_src else None
        self.pc = t(P.pc) if self.has_src else None
        self.pw = t(P.pw) if self.has_src else None
        self.g = S.Context.grid(P.nx, P.ny, P.nz, P.h, P.visc, P.bc_type)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle_solution(P, eps=1e-12):
    perm, indcol, b = O.assemble(P)
    cnt, col, ele = O.build_csr(perm, indcol)
    x, it = O.AMG(cnt, col, ele).pcg(b, eps=eps)
    return x, perm, indcol, b


def true_residual(perm, indcol, b, x):
    """||b - A x|| / ||b|| with A from the AoS stencil (numpy, FP64)."""
    Ax = O.prod_mat_vec(perm, indcol, x)
    return float(np.linalg.norm(b - Ax) / np.linalg.norm(b))
