"""ctypes access to the CPU restatement in oracle/ (the checker).

Only tests, the smoke check and the benchmark's CPU-baseline leg use this.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")

ELE = np.dtype([(f, "<f8") for f in ("u", "d", "l", "r", "c", "t", "b")])
IND = np.dtype([(f, "<i4") for f in ("c", "u", "d", "l", "r", "t", "b")])


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("hx", ctypes.c_double), ("hy", ctypes.c_double), ("hz", ctypes.c_double),
                ("visc", ctypes.c_double), ("bc_type", ctypes.c_int * 6)]


class Bound(ctypes.Structure):
    _fields_ = [(f, ctypes.c_void_p) for f in ("l", "r", "u", "d", "t", "b")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"], check=True)
        L = ctypes.CDLL(LIB)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.oracle_permeabilidade_arestas.argtypes = [vp] * 8
        L.oracle_source_term.argtypes = [vp] * 8
        L.oracle_build_csr.argtypes = [i, vp, vp, vp, vp, vp]
        L.oracle_build_csr.restype = ctypes.c_long
        L.oracle_prod_mat_vec.argtypes = [i, vp, vp, vp, vp]
        L.oracle_amg_create.argtypes = [i, vp, vp, vp, i, i, d]
        L.oracle_amg_create.restype = vp
        L.oracle_amg_destroy.argtypes = [vp]
        L.oracle_amg_levels.argtypes = [vp]
        L.oracle_amg_level_size.argtypes = [vp, i, vp, vp]
        L.oracle_amg_apply.argtypes = [vp, vp, vp]
        L.oracle_amg_level_matrix.argtypes = [vp, i, vp, vp, vp]
        L.oracle_amg_level_matrix.restype = ctypes.c_long
        L.oracle_amg_level_interp.argtypes = [vp, i, vp, vp, vp]
        L.oracle_amg_level_interp.restype = ctypes.c_long
        L.oracle_pcg.argtypes = [i, vp, vp, vp, vp, vp, vp, d, i]
        L.oracle_compute.argtypes = [i, i, vp, vp, vp, vp, vp, d]
        L.oracle_fluxos.argtypes = [vp] * 5
        L.oracle_divvel_pontual.argtypes = [vp] * 4
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data


class Problem:
    """A Darcy problem on the unit-cube-like box, as the reference's globals hold it."""

    def __init__(self, dims, K, bc_type, bc_val, visc=1.0, lengths=(1.0, 1.0, 1.0), KI=None, WI=None, pc=None, pw=None):
        self.nx, self.ny, self.nz = map(int, dims)
        self.n = self.nx * self.ny * self.nz
        n = self.n
        self.Kx = self.Ky = self.Kz = np.ascontiguousarray(K, np.float64)
        self.bc_type = tuple(int(t) for t in bc_type)
        self.bc_val = tuple(float(v) for v in bc_val)
        self.visc = float(visc)
        self.h = (lengths[0] / self.nx, lengths[1] / self.ny, lengths[2] / self.nz)
        z = np.zeros(n)
        self.KI = z if KI is None else np.ascontiguousarray(KI, np.float64)
        self.WI = z if WI is None else np.ascontiguousarray(WI, np.float64)
        self.pc = z if pc is None else np.ascontiguousarray(pc, np.float64)
        self.pw = z if pw is None else np.ascontiguousarray(pw, np.float64)
        # per-face value arrays, sized like the reference's Bound arrays (n each)
        self.B = [np.full(n, v) for v in self.bc_val]

    def grid(self):
        g = Grid(self.nx, self.ny, self.nz, self.h[0], self.h[1], self.h[2], self.visc)
        for f in range(6):
            g.bc_type[f] = self.bc_type[f]
        return g

    def bound(self):
        return Bound(*[_p(a) for a in self.B])


def assemble(P):
    g = P.grid()
    perm = np.zeros(P.n, ELE)
    indcol = np.zeros(P.n, IND)
    lib().oracle_permeabilidade_arestas(ctypes.byref(g), _p(P.Kx), _p(P.Ky), _p(P.Kz), _p(P.WI), _p(P.KI),
                                        _p(perm), _p(indcol))
    b = np.zeros(P.n)
    B = P.bound()
    lib().oracle_source_term(ctypes.byref(g), ctypes.byref(B), _p(perm), _p(P.pc), _p(P.pw), _p(P.KI), _p(P.WI), _p(b))
    return perm, indcol, b


def build_csr(perm, indcol):
    n = perm.size
    cnt = np.zeros(n, np.int32)
    col = np.zeros(7 * n, np.int32)
    ele = np.zeros(7 * n, np.float64)
    nnz = lib().oracle_build_csr(n, _p(perm), _p(indcol), _p(cnt), _p(col), _p(ele))
    return cnt, col[:nnz].copy(), ele[:nnz].copy()


def prod_mat_vec(perm, indcol, x):
    res = np.zeros(perm.size)
    lib().oracle_prod_mat_vec(perm.size, _p(perm), _p(indcol), _p(np.ascontiguousarray(x, np.float64)), _p(res))
    return res


class AMG:
    def __init__(self, cnt, col, ele, max_level=16, min_nodes=1, epsilon=0.0):
        self.cnt, self.col, self.ele = (np.ascontiguousarray(cnt, np.int32), np.ascontiguousarray(col, np.int32),
                                        np.ascontiguousarray(ele, np.float64))
        self.n = self.cnt.size
        self.h = lib().oracle_amg_create(self.n, _p(self.cnt), _p(self.col), _p(self.ele), max_level, min_nodes, epsilon)

    def levels(self):
        out = []
        for l in range(lib().oracle_amg_levels(self.h)):
            r = ctypes.c_int(); z = ctypes.c_long()
            lib().oracle_amg_level_size(self.h, l, ctypes.byref(r), ctypes.byref(z))
            out.append((r.value, z.value))
        return out

    def level_matrix(self, l):
        rows = self.levels()[l][0]
        nnz = lib().oracle_amg_level_matrix(self.h, l, None, None, None)
        cnt = np.zeros(rows, np.int32); col = np.zeros(nnz, np.int32); ele = np.zeros(nnz)
        lib().oracle_amg_level_matrix(self.h, l, _p(cnt), _p(col), _p(ele))
        return cnt, col, ele

    def level_interp(self, l):
        rows = self.levels()[l][0]
        m = lib().oracle_amg_level_interp(self.h, l, None, None, None)
        cnt = np.zeros(rows, np.int32); col = np.zeros(m, np.int32); ele = np.zeros(m)
        lib().oracle_amg_level_interp(self.h, l, _p(cnt), _p(col), _p(ele))
        return cnt, col, ele

    def apply(self, f):
        u = np.zeros(self.n)
        lib().oracle_amg_apply(self.h, _p(np.ascontiguousarray(f, np.float64)), _p(u))
        return u

    def pcg(self, b, x0=None, eps=1e-8, max_iterations=10000):
        x = np.zeros(self.n) if x0 is None else np.array(x0, np.float64)
        it = lib().oracle_pcg(self.n, _p(self.cnt), _p(self.col), _p(self.ele), self.h,
                              _p(np.ascontiguousarray(b, np.float64)), _p(x), eps, max_iterations)
        return x, it

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_amg_destroy(self.h)
            self.h = None


def compute(cnt, col, ele, b, x0=None, eps=1e-8):
    n = len(cnt)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    cnt = np.ascontiguousarray(cnt, np.int32); col = np.ascontiguousarray(col, np.int32)
    ele = np.ascontiguousarray(ele, np.float64); b = np.ascontiguousarray(b, np.float64)
    it = lib().oracle_compute(n, col.size, _p(cnt), _p(col), _p(ele), _p(b), _p(x), eps)
    return x, it


def fluxos(P, p, perm):
    v = np.zeros(P.n, ELE)
    g = P.grid(); B = P.bound()
    lib().oracle_fluxos(ctypes.byref(g), _p(np.ascontiguousarray(p, np.float64)), _p(perm), ctypes.byref(B), _p(v))
    return v


def divvel_pontual(P, v):
    g = P.grid()
    mn = ctypes.c_double(); mx = ctypes.c_double()
    lib().oracle_divvel_pontual(ctypes.byref(g), _p(v), ctypes.byref(mn), ctypes.byref(mx))
    return mn.value, mx.value


def problem_from_fixture(z):
    kw = {}
    if "KI" in z.files:
        kw = dict(KI=z["KI"], WI=z["WI"], pc=z["pc"], pw=z["pw"])
    return Problem(tuple(z["dims"]), z["K"], tuple(z["bc_type"]), tuple(z["bc_val"]), float(z["visc"]), **kw)
