"""Python host layer over libhmfem_gpu's C ABI (include/hmfem_gpu.h).

Two groups of functions, mirroring the library:

* the reference's own functions on the pressure path, same names and
  argument meaning, on host numpy arrays (permeabilidade_arestas,
  source_term, solve_linsys, compute, fluxos, divvel_pontual, prod_mat_vec);
* a device-resident `Context` whose methods take device pointers (for
  example torch CUDA tensors) and run the whole reservoir() step on the GPU.

There is no CPU fallback: every call goes through the HIP library and
raises if it is not built or the GPU call fails.
"""
import ctypes
import os

import numpy as np

from ._lib import load_gpu

ELE = np.dtype([(f, "<f8") for f in ("u", "d", "l", "r", "c", "t", "b")])
IND = np.dtype([(f, "<i4") for f in ("c", "u", "d", "l", "r", "t", "b")])

c_int, c_dbl, c_vp, c_long = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_long

PREC_FP64, PREC_FP32, PREC_FP16 = 0, 1, 2
STOP_RESIDUAL, STOP_PRECONDITIONED = 0, 1
SMOOTH_JACOBI, SMOOTH_BLOCK_GS = 0, 1
PRECOND_AMG, PRECOND_JACOBI = 0, 1
OK, NOT_CONVERGED = 0, 1
PROFILE_CLASSES = ("cg_spmv", "cg_update", "cg_dots", "presmooth", "restrict",
                   "postsmooth", "jacobi", "coarse")


class Dim(ctypes.Structure):      # structs.h:93-104
    _fields_ = [(f, c_int) for f in ("x", "y", "z", "xy", "t", "px", "pz", "py")]


class GridH(ctypes.Structure):    # structs.h:83-88 (`grid`)
    _fields_ = [("x", c_dbl), ("y", c_dbl), ("z", c_dbl)]


class FACE(ctypes.Structure):     # structs.h:50-59
    _fields_ = [(f, c_int) for f in ("u", "d", "l", "r", "t", "b")]


class Bound(ctypes.Structure):    # structs.h:27-36
    _fields_ = [(f, c_vp) for f in ("u", "d", "l", "r", "t", "b")]


class Perm(ctypes.Structure):     # structs.h:71-79
    _fields_ = [("X", c_vp), ("Y", c_vp), ("Z", c_vp), ("S", c_vp), ("permeability", c_int)]


class HmfemGrid(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int), ("hx", c_dbl), ("hy", c_dbl), ("hz", c_dbl),
                ("viscosidade", c_dbl), ("bc_type", c_int * 6)]


class Options(ctypes.Structure):
    _fields_ = [("rtol", c_dbl), ("max_iterations", c_int), ("precision", c_int), ("stop_rule", c_int),
                ("nu", c_int), ("omega", c_dbl), ("ocf", c_dbl), ("coarse_max", c_int), ("check_every", c_int),
                ("tail_max", c_int), ("tail_gamma", c_int), ("use_graph", c_int), ("tail_lds", c_int),
                ("smoother", c_int), ("gs_sweeps", c_int), ("autotune", c_int), ("deflate", c_int),
                ("deflate_ratio", c_dbl), ("precond", c_int), ("reuse_setup", c_int), ("semicoarsen", c_int),
                ("block_compact", c_int)]


class Status(ctypes.Structure):
    _fields_ = [("status", c_int), ("iterations", c_int), ("levels", c_int), ("rel_residual", c_dbl),
                ("precond_ratio", c_dbl), ("t_assemble_ms", c_dbl), ("t_setup_ms", c_dbl),
                ("t_solve_ms", c_dbl), ("t_flux_ms", c_dbl), ("div_max", c_dbl), ("div_min", c_dbl),
                ("solve_bytes", c_dbl), ("deflation_clusters", c_int), ("deflation_cells", c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BlockStatus(ctypes.Structure):
    _fields_ = [("status", c_int), ("max_iterations", c_int), ("t_assemble_ms", c_dbl), ("t_setup_ms", c_dbl),
                ("t_solve_ms", c_dbl), ("t_flux_ms", c_dbl), ("footprint_bytes", c_dbl),
                ("deflation_clusters", c_int), ("deflation_cells", c_int),
                ("block_iterations", c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchStatus(ctypes.Structure):
    _fields_ = [("status", c_int), ("max_iterations", c_int), ("t_assemble_ms", c_dbl), ("t_setup_ms", c_dbl),
                ("t_solve_ms", c_dbl), ("t_flux_ms", c_dbl), ("deflation_clusters", c_int),
                ("deflation_cells", c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ConduitSetC(ctypes.Structure):
    _fields_ = [("count", c_int), ("cells", c_vp), ("start", c_vp), ("axis", c_vp), ("Kc", c_vp), ("bc_l", c_vp),
                ("bc_r", c_vp), ("cont", c_vp), ("ccont", c_vp), ("karst_index", c_vp), ("area", c_vp),
                ("breccias", c_vp), ("max_breccias", c_int), ("breccia_cell", c_vp), ("breccia_beta", c_vp),
                ("breccia_C", c_vp), ("rho", c_dbl)]


class CoupledStatus(ctypes.Structure):
    _fields_ = [("status", c_int), ("passes", c_int), ("change", c_dbl), ("iterations", c_int),
                ("t_reservoir_ms", c_dbl), ("t_conduit_ms", c_dbl), ("t_total_ms", c_dbl)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ConduitSet:
    """A conduit set (the reference's data file, main.c:197-300) as host
    arrays; .struct() is the hmfem_conduit_set the library takes. Keyword
    arrays: cells, start (3 per conduit), axis, Kc, bc_l, bc_r, cont and ccont
    ([side * count + j]), karst_index, area, breccias; breccia_cell,
    breccia_beta, breccia_C as [max_breccias][count]; rho."""

    def __init__(self, cells, start, axis, Kc, bc_l, bc_r, cont, ccont, karst_index, area, breccias,
                 breccia_cell=None, breccia_beta=None, breccia_C=None, rho=1000.0):
        i32 = lambda a: np.ascontiguousarray(a, np.int32).reshape(-1)
        f64 = lambda a: np.ascontiguousarray(a, np.float64).reshape(-1)
        self.cells, self.start, self.axis = i32(cells), i32(start), i32(axis)
        self.count = int(self.cells.size)
        self.Kc, self.bc_l, self.bc_r = f64(Kc), i32(bc_l), i32(bc_r)
        self.cont, self.ccont = f64(cont), f64(ccont)
        self.karst_index, self.area, self.breccias = f64(karst_index), f64(area), i32(breccias)
        if breccia_cell is None or np.asarray(breccia_cell).size == 0:
            self.max_breccias = 0
            self.breccia_cell = np.zeros(1, np.int32)
            self.breccia_beta = np.zeros(1)
            self.breccia_C = np.zeros(1)
        else:
            self.max_breccias = int(np.asarray(breccia_cell).reshape(-1, self.count).shape[0])
            self.breccia_cell, self.breccia_beta, self.breccia_C = i32(breccia_cell), f64(breccia_beta), f64(breccia_C)
        self.rho = float(rho)
        self.total_cells = int(self.cells.sum())

    def struct(self):
        P = lambda a: a.ctypes.data
        return ConduitSetC(self.count, P(self.cells), P(self.start), P(self.axis), P(self.Kc), P(self.bc_l),
                           P(self.bc_r), P(self.cont), P(self.ccont), P(self.karst_index), P(self.area),
                           P(self.breccias), self.max_breccias, P(self.breccia_cell), P(self.breccia_beta),
                           P(self.breccia_C), self.rho)


_typed = False


def lib():
    global _typed
    L = load_gpu()
    if not _typed:
        P = ctypes.POINTER
        L.solve_linsys.argtypes = [c_vp, c_vp, c_vp, c_vp, Dim]
        L.solve_linsys.restype = None
        L.compute.argtypes = [c_vp] * 6
        L.permeabilidade_arestas.argtypes = [P(Perm), c_vp, c_vp, GridH, Dim, FACE, c_vp, c_vp, c_dbl]
        L.source_term.argtypes = [P(Bound), c_vp, GridH, Dim, FACE, c_vp, c_vp, c_vp, c_vp, c_dbl, c_vp]
        L.source_term.restype = None
        L.fluxos.argtypes = [c_vp, c_vp, c_vp, P(Bound), GridH, Dim]
        L.divvel_pontual.argtypes = [c_vp, GridH, Dim]
        L.divvel_pontual.restype = None
        L.prod_mat_vec.argtypes = [c_vp, c_vp, c_vp, c_vp, Dim]
        L.prod_mat_vec.restype = None
        L.hmfem_default_options.argtypes = [P(Options)]
        L.hmfem_default_options.restype = None
        L.hmfem_create.argtypes = [P(c_vp)]
        L.hmfem_destroy.argtypes = [c_vp]
        L.hmfem_destroy.restype = None
        L.hmfem_stream.argtypes = [c_vp]
        L.hmfem_stream.restype = c_vp
        L.hmfem_assemble_device.argtypes = [c_vp, P(HmfemGrid)] + [c_vp] * 7 + [c_vp, c_vp, c_vp, c_vp]
        L.hmfem_reservoir_device.argtypes = [c_vp, P(HmfemGrid)] + [c_vp] * 7 + [c_vp, c_vp, c_vp, P(Options),
                                                                              P(Status)]
        L.hmfem_flux_device.argtypes = [c_vp, P(HmfemGrid), c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, P(c_dbl), P(c_dbl)]
        L.hmfem_coupled_device.argtypes = [c_vp, P(HmfemGrid), c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, P(ConduitSetC),
                                           c_int, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, P(Options),
                                           P(CoupledStatus)]
        L.hmfem_hierarchy_level.argtypes = [c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]
        L.hmfem_hierarchy_coarse_inverse.argtypes = [c_vp, c_vp]
        L.hmfem_flux_device_noalloc.argtypes = [c_vp, P(HmfemGrid), c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, P(c_dbl),
                                                P(c_dbl)]
        L.hmfem_solve_structured_device.argtypes = [c_vp, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, P(Options),
                                                    P(Status)]
        L.hmfem_profile_enable.argtypes = [c_vp, c_int]
        L.hmfem_profile_read.argtypes = [c_vp, c_vp, c_vp, c_vp, c_int]
        L.hmfem_profile_reset.argtypes = [c_vp]
        L.hmfem_profile_reset.restype = None
        L.hmfem_reservoir_batch_device.argtypes = [c_vp, c_int, c_vp, c_vp, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp,
                                                   c_vp, c_vp, P(Options), P(BatchStatus), c_vp, c_vp, c_vp, c_vp]
        L.hmfem_reservoir_block_device.argtypes = [c_vp, P(HmfemGrid), c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp,
                                                   c_vp, P(Options), P(BlockStatus), c_vp, c_vp, c_vp]
        L.hmfem_autotune_device.argtypes = [c_vp, P(HmfemGrid), c_vp, c_vp, c_vp, c_vp, P(Options), P(c_dbl)]
        L.hmfem_autotune_batch_device.argtypes = [c_vp, c_int, c_vp, c_vp, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp,
                                                  c_vp, P(Options), c_vp, c_vp, P(c_dbl)]
        L.hmfem_kernel_bytes_per_cell.argtypes = [c_int, c_int]
        L.hmfem_kernel_bytes_per_cell.restype = c_dbl
        L.save_pressure.argtypes = [c_vp, Dim, GridH]
        L.save_pressure.restype = None
        L.save_flux.argtypes = [c_vp, Dim, GridH]
        L.save_flux.restype = None
        L.vts_StructuredGrid_file.argtypes = [c_vp, c_vp, Dim, GridH]
        L.save_pressure_1d.argtypes = [c_vp, c_vp, c_vp, GridH]
        L.save_pressure_1d.restype = None
        L.save_flux_1d.argtypes = [c_vp, c_vp, c_vp, GridH]
        L.save_flux_1d.restype = None
        L.hmfem_set_output_dir.argtypes = [ctypes.c_char_p]
        L.hmfem_set_output_dir.restype = None
        _typed = True
    return L


EXPORTED = ("solve_linsys", "compute", "permeabilidade_arestas", "source_term", "fluxos", "divvel_pontual",
            "prod_mat_vec", "hmfem_default_options", "hmfem_create", "hmfem_destroy", "hmfem_stream",
            "hmfem_assemble_device", "hmfem_reservoir_device", "hmfem_flux_device", "hmfem_flux_device_noalloc",
            "hmfem_coupled_device", "hmfem_hierarchy_level", "hmfem_hierarchy_coarse_inverse",
            "hmfem_solve_structured_device", "hmfem_profile_enable", "hmfem_profile_read",
            "hmfem_profile_reset", "hmfem_kernel_bytes_per_cell", "hmfem_reservoir_batch_device",
            "hmfem_reservoir_block_device", "hmfem_autotune_device", "hmfem_autotune_batch_device", "save_pressure",
            "save_flux",
            "vts_StructuredGrid_file", "hmfem_set_output_dir", "save_pressure_1d", "save_flux_1d", "iter_amg",
            "boundary", "dirname", "number_karst")


def _p(a):
    return a.ctypes.data


def make_dim(nx, ny, nz):
    return Dim(nx, ny, nz, nx * ny, nx * ny * nz, nx // 2, 0, ny // 2)


def face_types(bc_type):
    """(l, r, u, d, t, b) -> FACE"""
    l, r, u, d, t, b = (int(v) for v in bc_type)
    return FACE(u=u, d=d, l=l, r=r, t=t, b=b)


def default_options(**kw):
    o = Options()
    lib().hmfem_default_options(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


# ----------------------------------------------------------------- reference names


def iter_amg():
    return c_int.in_dll(lib(), "iter_amg").value


def set_boundary(bc_type):
    """Set the library's global `boundary` (read by fluxos, as in the reference)."""
    FACE.in_dll(lib(), "boundary").__init__(*[0] * 6)
    f = FACE.in_dll(lib(), "boundary")
    src = face_types(bc_type)
    for name, _ in FACE._fields_:
        setattr(f, name, getattr(src, name))


def permeabilidade_arestas(Kx, Ky, Kz, perm, indcol, h, N, bc_type, WI, KI, viscosidade):
    K = Perm(_p(Kx), _p(Ky), _p(Kz), None, 1)
    rc = lib().permeabilidade_arestas(ctypes.byref(K), _p(perm), _p(indcol), GridH(*h), N, face_types(bc_type),
                                      _p(WI), _p(KI), viscosidade)
    if rc != 0:
        raise RuntimeError("permeabilidade_arestas failed: %d" % rc)


def bound_struct(faces):
    """faces: six arrays in face order l, r, u, d, t, b."""
    l, r, u, d, t, b = (_p(a) for a in faces)
    return Bound(u=u, d=d, l=l, r=r, t=t, b=b)


def source_term(faces, perm, h, N, bc_type, pc, pw, KI, WI, viscosidade, b):
    B = bound_struct(faces)
    lib().source_term(ctypes.byref(B), _p(perm), GridH(*h), N, face_types(bc_type), _p(pc), _p(pw), _p(KI), _p(WI),
                      viscosidade, _p(b))


def solve_linsys(perm, indcol, b, x, N):
    lib().solve_linsys(_p(perm), _p(indcol), _p(b), _p(x), N)


def compute(hdr, cnt, col, ele, b, x):
    rc = lib().compute(_p(hdr), _p(cnt), _p(col), _p(ele), _p(b), _p(x))
    if rc != 0:
        raise RuntimeError("compute failed: %d" % rc)
    return rc


def fluxos(p, perm, v, faces, h, N):
    B = bound_struct(faces)
    return lib().fluxos(_p(p), _p(perm), _p(v), ctypes.byref(B), GridH(*h), N)


def divvel_pontual(v, h, N):
    lib().divvel_pontual(_p(v), GridH(*h), N)


def prod_mat_vec(perm, indcol, x, res, N):
    lib().prod_mat_vec(_p(perm), _p(indcol), _p(x), _p(res), N)


_dirname_keep = None


def set_dirname(path):
    """The output directory of the writers (hmfem_set_output_dir)."""
    global _dirname_keep
    _dirname_keep = ctypes.create_string_buffer(os.fsencode(path))
    lib().hmfem_set_output_dir(_dirname_keep)


def save_pressure(x, N, h):
    """pressure_reservoir.dat in the `dirname` directory (posprocessing_output.c:200)."""
    lib().save_pressure(_p(np.ascontiguousarray(x, np.float64)), N, GridH(*h))


def save_flux(v, N, h):
    """flux_reservoir.dat (posprocessing_output.c:236); v: ELE array (host)."""
    lib().save_flux(_p(v), N, GridH(*h))


class DimK(ctypes.Structure):     # structs.h:107-115
    _fields_ = [(f, c_int) for f in ("s", "sxi", "syi", "szi", "pos")]


class ELEK(ctypes.Structure):     # structs.h:19-25
    _fields_ = [("l", c_vp), ("r", c_vp), ("c", c_vp)]


def _rows(a):
    """double ** over the rows of a C-contiguous float64 matrix (kept alive by the caller)."""
    return (c_vp * a.shape[0])(*[a.ctypes.data + i * a.strides[0] for i in range(a.shape[0])])


def _conduit_dims(cells, start, axis):
    M = (DimK * len(cells))()
    for j, m in enumerate(cells):
        M[j] = DimK(int(m), int(start[3 * j]), int(start[3 * j + 1]), int(start[3 * j + 2]), int(axis[j]))
    c_int.in_dll(lib(), "number_karst").value = len(cells)
    return M


def save_pressure_1d(pc, cells, start, axis, h):
    """pressure_karst_<j>.dat (posprocessing_output.c:437); pc: [cell][conduit]."""
    x = np.ascontiguousarray(pc, np.float64)
    rows = _rows(x)
    M = _conduit_dims(cells, start, axis)
    hs = np.zeros(len(cells))
    lib().save_pressure_1d(rows, M, hs.ctypes.data, GridH(*h))


def save_flux_1d(ucl, ucr, cells, start, axis, h):
    """flux_karst_<j>.dat (posprocessing_output.c:536); ucl, ucr: [cell][conduit]."""
    l = np.ascontiguousarray(ucl, np.float64)
    r = np.ascontiguousarray(ucr, np.float64)
    rl, rr = _rows(l), _rows(r)
    uc = ELEK(ctypes.cast(rl, c_vp), ctypes.cast(rr, c_vp), None)
    M = _conduit_dims(cells, start, axis)
    hs = np.zeros(len(cells))
    lib().save_flux_1d(ctypes.byref(uc), M, hs.ctypes.data, GridH(*h))


def vts_StructuredGrid_file(v, x, N, h):
    """campos.vts (vtk.c:8); returns 0, or -1 if the file cannot be opened."""
    return lib().vts_StructuredGrid_file(_p(v), _p(np.ascontiguousarray(x, np.float64)), N, GridH(*h))


# ----------------------------------------------------------------- device API


def _dptr(t):
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _order_after_torch(*tensors):
    """The library runs on its own stream. Work that torch queued on its
    current stream for these tensors (fills, copies) must finish first."""
    for t in tensors:
        if hasattr(t, "data_ptr") and getattr(t, "is_cuda", False):
            import torch
            torch.cuda.current_stream(t.device).synchronize()
            return


class Context:
    """A HIP stream plus cached device buffers of the library."""

    def __init__(self):
        h = c_vp()
        rc = lib().hmfem_create(ctypes.byref(h))
        if rc != 0:
            raise RuntimeError("hmfem_create failed: %d (no HIP device?)" % rc)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().hmfem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return lib().hmfem_stream(self.h)

    @staticmethod
    def grid(nx, ny, nz, h=None, viscosidade=1.0, bc_type=(0, 0, 1, 1, 1, 1)):
        if h is None:
            h = (1.0 / nx, 1.0 / ny, 1.0 / nz)
        g = HmfemGrid(nx, ny, nz, h[0], h[1], h[2], viscosidade)
        for f in range(6):
            g.bc_type[f] = int(bc_type[f])
        return g

    @staticmethod
    def _faces(bval):
        arr = (c_vp * 6)(*[_dptr(t) for t in bval])
        return arr

    def assemble(self, g, Kx, Ky, Kz, bval, b, perm=None, indcol=None, KI=None, WI=None, pc=None, pw=None):
        _order_after_torch(Kx, Ky, Kz, b, perm, indcol, KI, WI, pc, pw, *bval)
        rc = lib().hmfem_assemble_device(self.h, ctypes.byref(g), _dptr(Kx), _dptr(Ky), _dptr(Kz), _dptr(KI),
                                         _dptr(WI), _dptr(pc), _dptr(pw), self._faces(bval), _dptr(perm),
                                         _dptr(indcol), _dptr(b))
        if rc != 0:
            raise RuntimeError("hmfem_assemble_device failed: %d" % rc)

    def reservoir(self, g, Kx, Ky, Kz, bval, p, v=None, options=None, KI=None, WI=None, pc=None, pw=None):
        st = Status()
        o = options if options is not None else default_options()
        _order_after_torch(Kx, Ky, Kz, p, v, KI, WI, pc, pw, *bval)
        rc = lib().hmfem_reservoir_device(self.h, ctypes.byref(g), _dptr(Kx), _dptr(Ky), _dptr(Kz), _dptr(KI),
                                          _dptr(WI), _dptr(pc), _dptr(pw), self._faces(bval), _dptr(p), _dptr(v),
                                          ctypes.byref(o), ctypes.byref(st))
        if rc < 0:
            raise RuntimeError("hmfem_reservoir_device failed: %d" % rc)
        return st

    def flux(self, g, Kx, Ky, Kz, bval, p, v=None):
        mn, mx = c_dbl(), c_dbl()
        _order_after_torch(Kx, Ky, Kz, p, v, *bval)
        rc = lib().hmfem_flux_device(self.h, ctypes.byref(g), _dptr(Kx), _dptr(Ky), _dptr(Kz), self._faces(bval),
                                     _dptr(p), _dptr(v), ctypes.byref(mn), ctypes.byref(mx))
        if rc != 0:
            raise RuntimeError("hmfem_flux_device failed: %d" % rc)
        return mn.value, mx.value

    def coupled(self, g, Kx, Ky, Kz, bval, conduits, p, pc, WI=None, pw=None, max_passes=1000, tol=1e-8,
                ucl=None, ucr=None, v=None, pc_first=None, options=None):
        """hmfem_coupled_device: main.c's reservoir / conduit fixed-point loop.
        Returns (status, per-pass relative changes)."""
        st = CoupledStatus()
        changes = np.zeros(max(max_passes, 1))
        cs = conduits.struct()
        _order_after_torch(Kx, Ky, Kz, p, pc, WI, pw, ucl, ucr, v, pc_first, *bval)
        rc = lib().hmfem_coupled_device(self.h, ctypes.byref(g), _dptr(Kx), _dptr(Ky), _dptr(Kz), self._faces(bval),
                                        _dptr(WI), _dptr(pw), ctypes.byref(cs), max_passes, tol, _dptr(p), _dptr(pc),
                                        _dptr(ucl), _dptr(ucr), _dptr(v), _dptr(pc_first), changes.ctypes.data,
                                        ctypes.byref(options) if options is not None else None, ctypes.byref(st))
        if rc < 0:
            raise RuntimeError("hmfem_coupled_device failed: %d" % rc)
        return st, changes[:st.passes]

    def hierarchy(self):
        """Levels of the last single AMG solve: [(dims, factors, e, wx, wy, wz)] (host, FP64)."""
        nlev = lib().hmfem_hierarchy_level(self.h, -1, None, None, None, None, None, None)
        if nlev < 0:
            raise RuntimeError("hmfem_hierarchy_level failed: %d" % nlev)
        out = []
        for l in range(nlev):
            d = np.zeros(3, np.int32)
            f = np.zeros(3, np.int32)
            lib().hmfem_hierarchy_level(self.h, l, d.ctypes.data, None, None, None, None, None)
            n = int(np.prod(d))
            e, wx, wy, wz = (np.zeros(n) for _ in range(4))
            rc = lib().hmfem_hierarchy_level(self.h, l, d.ctypes.data, f.ctypes.data, e.ctypes.data, wx.ctypes.data,
                                             wy.ctypes.data, wz.ctypes.data)
            if rc < 0:
                raise RuntimeError("hmfem_hierarchy_level failed: %d" % rc)
            out.append((tuple(int(v) for v in d), tuple(int(v) for v in f), e, wx, wy, wz))
        return out

    def coarse_inverse(self):
        """The dense inverse of the last hierarchy's coarsest level (host, nc x nc)."""
        nc = lib().hmfem_hierarchy_coarse_inverse(self.h, None)
        if nc < 0:
            raise RuntimeError("hmfem_hierarchy_coarse_inverse failed: %d" % nc)
        inv = np.zeros((nc, nc))
        lib().hmfem_hierarchy_coarse_inverse(self.h, inv.ctypes.data)
        return inv

    def flux_noalloc(self, g, Kx, Ky, Kz, bval, p, v=None):
        """hmfem_flux_device_noalloc: the same results, no device allocation."""
        mn, mx = c_dbl(), c_dbl()
        _order_after_torch(Kx, Ky, Kz, p, v, *bval)
        rc = lib().hmfem_flux_device_noalloc(self.h, ctypes.byref(g), _dptr(Kx), _dptr(Ky), _dptr(Kz),
                                             self._faces(bval), _dptr(p), _dptr(v), ctypes.byref(mn),
                                             ctypes.byref(mx))
        if rc != 0:
            raise RuntimeError("hmfem_flux_device_noalloc failed: %d" % rc)
        return mn.value, mx.value

    def solve_structured(self, nx, ny, nz, perm, indcol, b, x, options=None):
        st = Status()
        o = options if options is not None else default_options()
        _order_after_torch(perm, indcol, b, x)
        rc = lib().hmfem_solve_structured_device(self.h, nx, ny, nz, _dptr(perm), _dptr(indcol), _dptr(b), _dptr(x),
                                                 ctypes.byref(o), ctypes.byref(st))
        if rc < 0:
            raise RuntimeError("hmfem_solve_structured_device failed: %d" % rc)
        return st

    def reservoir_batch(self, dims, Kx, Ky, Kz, bval, p, v=None, options=None, viscosidade=1.0,
                        bc_type=(0, 0, 1, 1, 1, 1), lengths=None):
        """Solve len(dims) independent grids together. K*, p, v: concatenated
        device arrays; bval: six device arrays at least as long as the largest
        face. Returns (BatchStatus, per-instance dict of numpy arrays)."""
        dims_a = np.ascontiguousarray(np.asarray(dims, np.int32).reshape(-1, 3))
        ninst = dims_a.shape[0]
        len_a = None if lengths is None else np.ascontiguousarray(np.asarray(lengths, np.float64).reshape(-1, 3))
        bct = np.ascontiguousarray(np.asarray(bc_type, np.int32))
        its = np.zeros(ninst, np.int32)
        rr = np.zeros(ninst)
        dm = np.zeros(ninst)
        dv = np.zeros(ninst)
        st = BatchStatus()
        o = options if options is not None else default_options()
        _order_after_torch(Kx, Ky, Kz, p, v, *bval)
        rc = lib().hmfem_reservoir_batch_device(self.h, ninst, _p(dims_a), None if len_a is None else _p(len_a),
                                                viscosidade, _p(bct), self._faces(bval), _dptr(Kx), _dptr(Ky),
                                                _dptr(Kz), _dptr(p), _dptr(v), ctypes.byref(o), ctypes.byref(st),
                                                _p(its), _p(rr), _p(dm), _p(dv))
        if rc < 0:
            raise RuntimeError("hmfem_reservoir_batch_device failed: %d" % rc)
        return st, dict(iterations=its, rel_residual=rr, done_ms=dm, div_max=dv)

    def autotune(self, g, Kx, Ky, Kz, bval, options=None):
        """Choose cycle options for this problem; returns (Options, best ms)."""
        o = Options()
        ctypes.memmove(ctypes.byref(o), ctypes.byref(options if options is not None else default_options()),
                       ctypes.sizeof(Options))
        ms = c_dbl()
        _order_after_torch(Kx, Ky, Kz, *bval)
        rc = lib().hmfem_autotune_device(self.h, ctypes.byref(g), _dptr(Kx), _dptr(Ky), _dptr(Kz), self._faces(bval),
                                         ctypes.byref(o), ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError("hmfem_autotune_device failed: %d" % rc)
        return o, ms.value

    def autotune_batch(self, dims, Kx, Ky, Kz, bval, p=None, options=None, viscosidade=1.0,
                       bc_type=(0, 0, 1, 1, 1, 1), lengths=None):
        """Per-instance block Gauss-Seidel sweeps and coarse-correction scaling
        for this batch (p: the initial guess, kept; None: zero); returns
        (sweeps, scaling, best ms) with numpy arrays of one entry per instance."""
        dims_a = np.ascontiguousarray(np.asarray(dims, np.int32).reshape(-1, 3))
        ninst = dims_a.shape[0]
        len_a = None if lengths is None else np.ascontiguousarray(np.asarray(lengths, np.float64).reshape(-1, 3))
        bct = np.ascontiguousarray(np.asarray(bc_type, np.int32))
        gs = np.zeros(ninst, np.int32)
        ocf = np.zeros(ninst)
        ms = c_dbl()
        o = options if options is not None else default_options()
        _order_after_torch(Kx, Ky, Kz, p, *bval)
        rc = lib().hmfem_autotune_batch_device(self.h, ninst, _p(dims_a), None if len_a is None else _p(len_a),
                                               viscosidade, _p(bct), self._faces(bval), _dptr(Kx), _dptr(Ky),
                                               _dptr(Kz), _dptr(p), ctypes.byref(o), _p(gs), _p(ocf),
                                               ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError("hmfem_autotune_batch_device failed: %d" % rc)
        return gs, ocf, ms.value

    def reservoir_block(self, g, Kx, Ky, Kz, bvals, p, v=None, q=None, options=None):
        """Solve len(p) right-hand sides of one grid together. bvals[r]: six
        device face arrays of right-hand side r; p[r]: device pressure (initial
        guess in, solution out); v[r]: device ELE fluxes or None; q[r]: device
        array added to b, or None. Returns (BlockStatus, per-RHS dict)."""
        nr = len(p)
        fa = (c_vp * (6 * nr))(*[_dptr(t) for bv in bvals for t in bv])
        pa = (c_vp * nr)(*[_dptr(t) for t in p])
        va = None if v is None else (c_vp * nr)(*[_dptr(t) for t in v])
        qa = None if q is None else (c_vp * nr)(*[_dptr(t) for t in q])
        its = np.zeros(nr, np.int32)
        rr = np.zeros(nr)
        dv = np.zeros(nr)
        st = BlockStatus()
        o = options if options is not None else default_options()
        _order_after_torch(Kx, Ky, Kz, *p, *(v or []), *(t for t in (q or []) if t is not None),
                           *[t for bv in bvals for t in bv])
        rc = lib().hmfem_reservoir_block_device(self.h, ctypes.byref(g), _dptr(Kx), _dptr(Ky), _dptr(Kz), nr, fa,
                                                qa, pa, va, ctypes.byref(o), ctypes.byref(st), _p(its), _p(rr),
                                                _p(dv))
        if rc < 0:
            raise RuntimeError("hmfem_reservoir_block_device failed: %d" % rc)
        return st, dict(iterations=its, rel_residual=rr, div_max=dv)

    def profile_enable(self, max_launches=20000):
        rc = lib().hmfem_profile_enable(self.h, max_launches)
        if rc != 0:
            raise RuntimeError("hmfem_profile_enable failed: %d" % rc)

    def profile_reset(self):
        lib().hmfem_profile_reset(self.h)

    def profile_read(self):
        k = len(PROFILE_CLASSES)
        ms = np.zeros(k)
        cnt = np.zeros(k, np.int64)
        by = np.zeros(k)
        lib().hmfem_profile_read(self.h, _p(ms), _p(cnt), _p(by), k)
        return {name: dict(ms=float(ms[i]), count=int(cnt[i]), bytes_per_launch=float(by[i]))
                for i, name in enumerate(PROFILE_CLASSES)}
