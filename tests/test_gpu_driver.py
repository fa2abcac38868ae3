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
