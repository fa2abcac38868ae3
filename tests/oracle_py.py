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
        L.oracle_conduit.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, d, d, i]
        L.oracle_conduit.restype = None
        L.oracle_fluxo1d.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, d, vp, vp, vp]
        L.oracle_fluxo1d.restype = None
        L.oracle_rel_change.argtypes = [i, vp, vp]
        L.oracle_rel_change.restype = d
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


# ------------------------------------------------------------------ 1-D conduits


class _Conduits(ctypes.Structure):
    _fields_ = [("count", ctypes.c_int), ("mmax", ctypes.c_int), ("max_brec", ctypes.c_int)] + \
        [(f, ctypes.c_void_p) for f in ("cells", "sx", "sy", "sz", "axis", "bc_l", "bc_r", "cont", "ccont", "nbrec",
                                        "brec", "beta", "C")] + [("rho", ctypes.c_double)]


class _Karst(ctypes.Structure):
    _fields_ = [(f, ctypes.c_void_p) for f in ("KS", "KI_c", "hs", "p_aux", "pc", "pc_old", "ucR_old", "ucL_old",
                                               "kl", "kr", "kc", "ind_l", "ind_c", "ind_r")]


class ConduitSet:
    """A conduit set as main.c:197-300 holds it (arrays [row][conduit])."""

    def __init__(self, z):
        i32 = lambda a: np.ascontiguousarray(a, np.int32)
        f64 = lambda a: np.ascontiguousarray(a, np.float64)
        self.cells = i32(z["cells"])
        self.count = int(self.cells.size)
        st = i32(z["start"]).reshape(self.count, 3)
        self.start = i32(z["start"])
        self.sx, self.sy, self.sz = i32(st[:, 0]), i32(st[:, 1]), i32(st[:, 2])
        self.axis = i32(z["axis"])
        bc = i32(z["bc"]).reshape(self.count, 2)
        self.bc_l, self.bc_r = i32(bc[:, 0]), i32(bc[:, 1])
        self.Kc = f64(z["Kc"])
        self.cont = f64(z["cont"])           # [side * count + j]
        self.ccont = f64(z["ccont"])
        self.karst_index = f64(z["karst_index"])
        self.area = f64(z["area"])
        self.nbrec = i32(z["nbrec"])
        self.brec = i32(z["brec"]).reshape(-1)
        self.beta = f64(z["beta"]).reshape(-1)
        self.C = f64(z["C"]).reshape(-1)
        self.max_brec = int(np.asarray(z["brec"]).shape[0])
        self.mmax = int(self.cells.max())
        self.rho = float(z["rho"])

    def struct(self):
        return _Conduits(self.count, self.mmax, self.max_brec, *[_p(getattr(self, f)) for f in (
            "cells", "sx", "sy", "sz", "axis", "bc_l", "bc_r", "cont", "ccont", "nbrec", "brec", "beta", "C")],
            self.rho)

    def reservoir_cells(self, dims):
        """Reservoir index of conduit j's cell i, as [(j, i, s)] in main.c's loop order."""
        nx, ny, nz = dims
        out = []
        for j in range(self.count):
            for i in range(int(self.cells[j])):
                x, y, zz = int(self.sx[j]), int(self.sy[j]), int(self.sz[j])
                if self.axis[j] == 1:
                    x += i
                elif self.axis[j] == 2:
                    y += i
                else:
                    zz += i
                out.append((j, i, x + nx * (y + ny * zz)))
        return out


def karst_reservoir(z):
    """The reservoir of main.c's setup: K, faces, well arrays."""
    dims = tuple(int(v) for v in z["dims"])
    nx, ny, nz = dims
    n = nx * ny * nz
    WI = np.zeros(n)
    pw = np.zeros(n)
    for k in range(int(z["well_z0"]), nz):
        s = nx // 2 + nx * (ny // 2 + ny * k)
        WI[s] = float(z["well_index"])
        pw[s] = float(z["ppoco"])
    P = Problem(dims, np.full(n, float(z["K"])), tuple(z["bc_type"]), tuple(z["bc_val"]), float(z["visc"]),
                tuple(float(v) for v in z["lengths"]), KI=np.zeros(n), WI=np.zeros(n), pc=np.zeros(n), pw=pw)
    return P, WI


def coupled(z, max_passes=1000, eps3d=1e-8, eps1d=1e-8, direct=False, tol=1e-8):
    """main.c:395-530 over the oracle: the reservoir pass without well and
    conduits, the conduit set's karst index, conduit(), then the fixed-point
    loop of reservoir() and conduit() until the reservoir pressure changes by
    at most `tol` between passes (main.c:457-514; the loop's max search keeps
    error1[number_karst], the reservoir change), then fluxo1d() and fluxos().
    eps3d / eps1d: compute()'s stop for the 3-D and 1-D systems (the
    reference: 1e-8); direct=True solves the 1-D systems exactly."""
    P, WI_well = karst_reservoir(z)
    cs = ConduitSet(z)
    nk, mm, mb = cs.count, cs.mmax, max(cs.max_brec, 1)
    g = P.grid()
    dims = (P.nx, P.ny, P.nz)
    hsz = {1: P.h[0], 2: P.h[1], 3: P.h[2]}
    KS = np.repeat(cs.Kc[None, :], mm, axis=0).copy()
    KI_c = np.zeros((mm, nk))
    hs = np.array([hsz[int(a)] for a in cs.axis])
    st = dict(KS=KS, KI_c=KI_c, hs=hs, p_aux=np.zeros((mm, nk)), pc=np.zeros((mm, nk)), pc_old=np.zeros((mm, nk)),
              ucR_old=np.zeros((mb, nk)), ucL_old=np.zeros((mb, nk)), kl=np.zeros((mm, nk)), kr=np.zeros((mm, nk)),
              kc=np.zeros((mm, nk)), ind_l=np.zeros(mm + 1, np.int32), ind_c=np.zeros(mm + 1, np.int32),
              ind_r=np.zeros(mm + 1, np.int32))
    K = _Karst(*[_p(st[f]) for f in ("KS", "KI_c", "hs", "p_aux", "pc", "pc_old", "ucR_old", "ucL_old", "kl", "kr",
                                      "kc", "ind_l", "ind_c", "ind_r")])
    C = cs.struct()
    cells = cs.reservoir_cells(dims)
    n = P.n
    p = np.zeros(n)
    iters = []

    def reservoir():
        pc_aux = np.zeros(n)
        for j, i, s in cells:
            pc_aux[s] = st["pc_old"][i, j]
        P.pc = pc_aux
        perm, indcol, b = assemble(P)
        cnt, col, ele = build_csr(perm, indcol)
        x, it = compute(cnt, col, ele, b, x0=p, eps=eps3d)
        p[:] = x
        iters.append(it)
        return perm

    def conduit(perm, p_old):
        lib().oracle_conduit(ctypes.byref(C), ctypes.byref(K), ctypes.byref(g), _p(P.Kx), _p(P.Ky), _p(P.Kz), _p(perm),
                             _p(p_old), P.visc, eps1d, 1 if direct else 0)

    perm = reservoir()
    p_old = p.copy()
    P.WI = WI_well.copy()
    KI = np.zeros(n)
    for j, i, s in cells:
        KI[s] = cs.karst_index[j]
        KI_c[i, j] = cs.karst_index[j] / (cs.area[j] * hs[j])
        st["p_aux"][i, j] = p_old[s]
    P.KI = KI
    conduit(perm, p_old)
    out = dict(pc0=st["pc"].copy())
    error, passes, changes = 1.0, 0, []
    while passes < max_passes and error > tol:
        st["pc_old"][:] = st["pc"]
        perm = reservoir()
        error = lib().oracle_rel_change(n, _p(p), _p(p_old))
        changes.append(error)
        p_old[:] = p
        for j, i, s in cells:
            st["p_aux"][i, j] = p_old[s]
        conduit(perm, p_old)
        passes += 1
        if passes == 1:
            out.update(p1=p.copy(), pc1=st["pc"].copy())
    ucl = np.zeros((mm, nk))
    ucr = np.zeros((mm, nk))
    lib().oracle_fluxo1d(ctypes.byref(C), ctypes.byref(K), ctypes.byref(g), _p(P.Kx), _p(P.Ky), _p(P.Kz), _p(perm),
                         _p(p_old), P.visc, _p(st["pc"]), _p(ucl), _p(ucr))
    v = fluxos(P, p, perm)
    out.update(p=p.copy(), pc=st["pc"].copy(), ucl=ucl, ucr=ucr, ucR_old=st["ucR_old"].copy(),
               ucL_old=st["ucL_old"].copy(), kc=st["kc"].copy(), v=v, passes=passes, error=error,
               pass_change=np.array(changes), iter_amg=np.array(iters, np.int32))
    return out
