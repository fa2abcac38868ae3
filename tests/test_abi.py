"""The C-ABI libraries load and export every function include/*.h declares
(no GPU needed: nothing here calls into the device)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"typedef[^;]*;", "", text)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(([^;{}]*)\)\s*;", text):
        name = m.group(1)
        if name not in ("if", "for", "while", "return", "sizeof"):
            names.add(name)
    return names


def declared_globals(header):
    text = open(header).read()
    return set(re.findall(r"^extern\s+\w+\s+(\w+)\s*;", text, flags=re.M))


@pytest.mark.parametrize("header,lib", [("hmfem_gpu.h", "libhmfem_gpu.so"), ("hmfem_fields.h", "libhmfem_fields.so")])
def test_exports(header, lib):
    path = os.path.join(ROOT, "hmfem_amd", lib)
    if not os.path.exists(path):
        pytest.fail("%s not built; run __graft_entry__.build()" % path)
    L = ctypes.CDLL(path)
    h = os.path.join(ROOT, "include", header)
    funcs = declared_functions(h)
    assert len(funcs) >= 4
    missing = [f for f in sorted(funcs) if not hasattr(L, f)]
    assert not missing, missing
    for g in declared_globals(h):
        ctypes.c_int.in_dll(L, g)


def test_reference_entry_points_declared():
    funcs = declared_functions(os.path.join(ROOT, "include", "hmfem_gpu.h"))
    for name in ("solve_linsys", "compute", "permeabilidade_arestas", "source_term", "fluxos", "divvel_pontual",
                 "prod_mat_vec", "hmfem_reservoir_device"):
        assert name in funcs


def test_python_layer_binds_everything():
    from hmfem_amd import solver as S
    funcs = declared_functions(os.path.join(ROOT, "include", "hmfem_gpu.h"))
    L = S.lib()
    for f in funcs:
        assert f in S.EXPORTED, f
        assert getattr(L, f).argtypes is not None or f in ("hmfem_default_options",), f


def test_options_defaults():
    from hmfem_amd import solver as S
    o = S.default_options()
    assert o.rtol == 1e-8 and o.max_iterations == 10000 and o.precision == S.PREC_FP32
    assert o.nu >= 1 and 0 < o.omega < 1 and o.coarse_max >= 1
    assert o.smoother in (S.SMOOTH_JACOBI, S.SMOOTH_BLOCK_GS) and o.gs_sweeps >= 0 and o.autotune == 0
    assert o.deflate == 1 and o.deflate_ratio == 3000.0 and o.precond == S.PRECOND_AMG and o.reuse_setup == 0


def test_compute_unsupported_matrix_ends_the_process():
    """compute() on a matrix it cannot solve (here: not symmetric) ends the
    process with a message rather than returning with x unchanged: the
    reference's callers ignore compute()'s return value
    (solve_linearsystem_AMG.c:150, :283). The check runs before any device
    work, so this needs no GPU."""
    import subprocess
    import sys
    code = r'''
import sys
import numpy as np
sys.path.insert(0, %r)
from hmfem_amd import solver as S
n = 4
cnt = np.array([2, 3, 3, 2], np.int32)
col = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3], np.int32)
ele = np.array([2.0, -1.0, -1.0, 2.0, -1.0, -0.5, 2.0, -1.0, -1.0, 2.0])   # a(2,1) != a(1,2)
x = np.zeros(n)
S.compute(np.array([n, col.size], np.int32), cnt, col, ele, np.ones(n), x)
print("RETURNED", x.tolist())
''' % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "RETURNED" not in r.stdout
    assert "compute" in r.stderr
