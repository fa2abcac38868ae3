"""The C host driver (hmfem_amd/hmfem_run) end to end on the GPU: three
reservoir passes with a growing well index, warm starts and, with
HMFEM_REUSE_SETUP=1, the first pass's hierarchy reused by the later ones; the
pressure file it writes holds the oracle's solution of the last pass's
system (to the file's 10 significant digits)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
from gpu_helpers import oracle_solution, rel_l2
from hmfem_amd import fields as F

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hmfem_amd", "hmfem_run")

CONFIG = """# 24 x 20 x 16 lognormal field with two conduits and one well
grid 24 20 16
bc_type 0 0 1 1 1 1
bc_value 1 0 0 0 0 0
field lognormal -32.236191301916641 1.5 4 7
conduits 2 9
well 12 10 8 1e-12 2.0
ki_growth 1.5
passes 3
"""


def read_pressure(path, n):
    vals = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 4:
                vals.append(float(parts[3]))
    assert len(vals) == n
    return np.array(vals)


@pytest.mark.parametrize("reuse", ["0", "1"])
def test_driver_passes(reuse, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("hmfem_run is not built")
    cfg = tmp_path / "cfg.txt"
    cfg.write_text(CONFIG)
    out = tmp_path / "out"
    out.mkdir()
    env = dict(os.environ, HMFEM_REUSE_SETUP=reuse, HMFEM_RTOL="1e-12")
    print("[driver] reuse", reuse, flush=True)
    r = subprocess.run([EXE, str(cfg), str(out)], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    its = [int(l.split()[-1]) for l in r.stdout.splitlines() if l.startswith("pass ")]
    assert len(its) == 3 and all(i >= 1 for i in its), r.stdout
    nx, ny, nz = 24, 20, 16
    n = nx * ny * nz
    K = F.conduits(F.lognormal(nx, ny, nz, -32.236191301916641, 1.5, 4.0, 7), nx, ny, nz, 2, 9)
    KI = np.zeros(n)
    pc = np.zeros(n)
    s = 12 + 10 * nx + 8 * nx * ny
    KI[s] = 1e-12 * 1.5 ** 2   # the last pass
    pc[s] = 2.0
    P = O.Problem((nx, ny, nz), K, (0, 0, 1, 1, 1, 1), (1.0, 0, 0, 0, 0, 0), KI=KI, pc=pc)
    xo, _, _, _ = oracle_solution(P, eps=1e-13)
    # pressure_reservoir.dat lists cells with k innermost (i, j, k loops)
    x = read_pressure(out / "pressure_reservoir.dat", n).reshape(nx, ny, nz).transpose(2, 1, 0).ravel()
    assert rel_l2(x, xo) < 1e-8
    for f in ("flux_reservoir.dat", "campos.vts"):
        assert (out / f).stat().st_size > 0


def write_conduit_file(z, path):
    """The fixture's conduit set in the reference's data-file format
    (main.c:197-283 reads it; beta holds R^4 - 1, the file R^4)."""
    nk = len(z["cells"])
    st = np.asarray(z["start"]).reshape(nk, 3)
    bc = np.asarray(z["bc"]).reshape(nk, 2)
    row = lambda vals: " ".join(repr(float(v)) if isinstance(v, (float, np.floating)) else str(int(v)) for v in vals)
    lines = [row(z["cells"]), row(st[:, 0]), row(st[:, 1]), row(st[:, 2]), row(z["axis"]),
             row(np.asarray(z["Kc"], float)), row(bc[:, 0]), row(bc[:, 1])]
    cont = np.asarray(z["cont"], float).reshape(2, nk)
    ccont = np.asarray(z["ccont"], float).reshape(2, nk)
    lines += [row(cont[0]), row(cont[1]), row(ccont[0]), row(ccont[1]), row(np.asarray(z["karst_index"], float)),
              row(np.asarray(z["area"], float)), row(z["nbrec"])]
    beta = np.asarray(z["beta"], float).reshape(-1, nk)
    lines += [row(b + 1.0) for b in beta]
    lines += [row(c) for c in np.asarray(z["C"], float).reshape(-1, nk)]
    lines += [row(b) for b in np.asarray(z["brec"]).reshape(-1, nk)]
    path.write_text("\n".join(lines) + "\n")


def read_columns(path, col):
    return np.array([float(l.split()[col]) for l in open(path) if len(l.split()) > col])


@pytest.mark.parametrize("case", ["ref_210x42x20_1_2", "mixed_48x40x30"])
def test_driver_coupled_conduits(case, tmp_path):
    """hmfem_run with a conduit data file runs main.c's coupled loop on the
    device and writes the reference's files; the conduit pressures it prints
    (pressure_karst_<j>.dat) agree with the reference's coupled run, and the
    pass count with the reference's."""
    if not os.path.exists(EXE):
        pytest.fail("hmfem_run is not built")
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "karst", case + ".npz"))
    nx, ny, nz = (int(v) for v in z["dims"])
    data = tmp_path / "conduits.dat"
    write_conduit_file(z, data)
    L = [float(v) for v in z["lengths"]]
    cfg = tmp_path / "cfg.txt"
    cfg.write_text("grid %d %d %d\nlengths %r %r %r\nviscosity %r\nbc_type %s\nbc_value %s\nfield homogeneous %r\n"
                   "well_column %d %d %d %r %r\nkarst_data %d %s\n"
                   % (nx, ny, nz, L[0], L[1], L[2], float(z["visc"]), " ".join(str(int(v)) for v in z["bc_type"]),
                      " ".join(repr(float(v)) for v in z["bc_val"]), float(z["K"]), nx // 2, ny // 2,
                      int(z["well_z0"]), float(z["well_index"]), float(z["ppoco"]), len(z["cells"]), data))
    out = tmp_path / "out"
    out.mkdir()
    print("[driver] coupled", case, flush=True)
    r = subprocess.run([EXE, str(cfg), str(out)], capture_output=True, text=True, timeout=240)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stderr
    passes = int([l for l in r.stdout.splitlines() if l.startswith("iterations number=")][0].split("=")[1])
    assert abs(passes - int(z["passes"])) <= max(3, int(z["passes"]) // 10)
    for j, m in enumerate(z["cells"]):
        pc = read_columns(out / ("pressure_karst_%d.dat" % j), 3)
        ref = z["pc"][:m, j]
        assert pc.size == m
        assert np.abs(pc - ref).max() <= 1e-7 * np.abs(ref).max(), j
        assert (out / ("flux_karst_%d.dat" % j)).stat().st_size > 0
    p = read_pressure(out / "pressure_reservoir.dat", nx * ny * nz)
    assert abs(np.linalg.norm(p) - float(z["p_norm2"])) <= 1e-7 * float(z["p_norm2"])
