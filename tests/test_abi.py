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
    assert o.smoother in (S.SMOOTH_JACOBI, S.SMOOTH_BLOCK_GS) and o.gs_sweeps >= 1 and o.autotune == 0
    assert o.deflate == 1 and o.deflate_ratio == 3000.0 and o.precond == S.PRECOND_AMG and o.reuse_setup == 0
