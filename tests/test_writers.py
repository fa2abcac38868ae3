"""Output writers against the reference's own files (CPU).

tests/golden/writers/*.npz hold, for two problems, the bytes that the
reference build's save_pressure(), save_flux() (posprocessing_output.c:200-296)
and vts_StructuredGrid_file() (vtk.c:8-76) wrote for the tight pressure p12
and its fluxes v; the library's writers must write the same bytes from the
same arrays.
"""
import glob
import os

import numpy as np
import pytest

from hmfem_amd import solver as S

ALL = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "writers", "*.npz")))
W = [p for p in ALL if not os.path.basename(p).startswith("conduits_")]
WK = [p for p in ALL if os.path.basename(p).startswith("conduits_")]


@pytest.mark.parametrize("path", W, ids=[os.path.basename(p)[:-4] for p in W])
def test_writers_bytes(path, tmp_path):
    z = np.load(path)
    nx, ny, nz = (int(a) for a in z["dims"])
    N = S.make_dim(nx, ny, nz)
    h = (1.0 / nx, 1.0 / ny, 1.0 / nz)
    v = np.zeros(nx * ny * nz, S.ELE)
    for k, f in enumerate(("u", "d", "l", "r", "c", "t", "b")):
        v[f] = z["v"][:, k]
    S.set_dirname(str(tmp_path))
    S.save_pressure(z["p12"], N, h)
    S.save_flux(v, N, h)
    assert S.vts_StructuredGrid_file(v, z["p12"], N, h) == 0
    for fname, key in (("pressure_reservoir.dat", "pressure_reservoir"), ("flux_reservoir.dat", "flux_reservoir"),
                       ("campos.vts", "campos")):
        got = (tmp_path / fname).read_bytes()
        ref = z[key].tobytes()
        assert len(got) == len(ref) and got == ref, fname


def test_writers_unwritable_dir(tmp_path):
    """campos.vts into a missing directory: -1 (the reference would write to a
    null stream)."""
    N = S.make_dim(2, 2, 2)
    v = np.zeros(8, S.ELE)
    S.set_dirname(str(tmp_path / "missing" / "dir"))
    assert S.vts_StructuredGrid_file(v, np.zeros(8), N, (0.5, 0.5, 0.5)) == -1


def test_writers_read_host_dirname(tmp_path):
    """A C host that defines `dirname` (as the reference's main.c does) gets
    the files in that directory: the library reads the executable's global."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "hmfem_amd")
    src = tmp_path / "host.c"
    out = tmp_path / "out"
    out.mkdir()
    src.write_text(
        '#include "hmfem_gpu.h"\n'
        "const char *dirname;\n"
        "int main(int argc, char **argv) {\n"
        "    double p[8] = {1, 2, 3, 4, 5, 6, 7, 8};\n"
        "    Dim N = {2, 2, 2, 4, 8, 0, 0, 0};\n"
        "    grid h = {0.5, 0.5, 0.5};\n"
        "    dirname = argv[1];\n"
        "    save_pressure(p, N, h);\n"
        "    return 0;\n"
        "}\n")
    exe = tmp_path / "host"
    subprocess.run(["gcc", "-o", str(exe), str(src), "-I" + os.path.join(root, "include"), "-L" + libdir,
                    "-lhmfem_gpu", "-Wl,-rpath," + libdir], check=True)
    subprocess.run([str(exe), str(out)], check=True, timeout=120)
    text = (out / "pressure_reservoir.dat").read_text()
    assert text.startswith("0.250000 0.250000 0.250000 1.000000000e+00\n")
    assert text.count("\n") == 8 + 4 + 2


def test_writers_long_lines(tmp_path):
    """Coordinates far beyond the usual unit box give lines longer than the
    formatter's line buffer; they must come out whole, as fprintf prints them."""
    N = S.make_dim(2, 1, 1)
    h = (1e300, 1.0, 1.0)
    S.set_dirname(str(tmp_path))
    S.save_pressure(np.array([1.0, 2.0]), N, h)
    text = (tmp_path / "pressure_reservoir.dat").read_text()
    first = text.splitlines()[0].split()
    assert len(first) == 4 and float(first[0]) == 0.5e300 and first[3] == "1.000000000e+00"
    assert first[0] == "%.6f" % 0.5e300


@pytest.mark.parametrize("path", WK, ids=[os.path.basename(p)[:-4] for p in WK])
def test_conduit_writers_bytes(path, tmp_path):
    """save_pressure_1d() and save_flux_1d() (posprocessing_output.c:437-659)
    against the files the reference's own writers wrote for the final
    conduit pressures and fluxes of a coupled run (every axis)."""
    z = np.load(path)
    dims = [int(a) for a in z["dims"]]
    h = tuple(float(L) / n for L, n in zip(z["lengths"], dims))
    S.set_dirname(str(tmp_path))
    S.save_pressure_1d(z["pc"], z["cells"], z["start"], z["axis"], h)
    S.save_flux_1d(z["ucl"], z["ucr"], z["cells"], z["start"], z["axis"], h)
    for j in range(len(z["cells"])):
        for stem in ("pressure_karst", "flux_karst"):
            got = (tmp_path / ("%s_%d.dat" % (stem, j))).read_bytes()
            ref = z["%s_%d" % (stem, j)].tobytes()
            assert got == ref, (stem, j)
