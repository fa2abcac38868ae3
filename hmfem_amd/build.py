"""Build the package's native libraries in-tree.

libhmfem_fields.so  host C (gcc): seeded permeability fields
libhmfem_gpu.so     HIP for gfx950 (hipcc): the solver and its C ABI
"""
import concurrent.futures
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")

HIP_SOURCES = ["hmfem_assemble.hip", "hmfem_solver.hip", "hmfem_batch.hip", "hmfem_block.hip", "hmfem_defl.hip",
               "hmfem_conduit.hip", "hmfem_csr.hip", "hmfem_capi.hip", "hmfem_output.cpp"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 arithmetic instructions: some of their forms have given
    # wrong results on this hardware without any error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    if verbose and r.stderr.strip():
        print(r.stderr[-4000:])


def build_fields(verbose=False):
    out = os.path.join(HERE, "libhmfem_fields.so")
    _run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-o", out,
          os.path.join(CSRC, "hmfem_fields.c"), "-lm"], verbose)
    return out


def build_gpu(verbose=False):
    """Compile every HIP source to an object in parallel, then link."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    out = os.path.join(HERE, "libhmfem_gpu.so")
    tmpdir = tempfile.mkdtemp(prefix="hmfem_build_")
    try:
        objs = [os.path.join(tmpdir, os.path.splitext(src)[0] + ".o") for src in HIP_SOURCES]
        cmds = [[hipcc] + HIPCC_FLAGS + ["-I" + os.path.join(ROOT, "include"), "-c", "-o", obj,
                                         os.path.join(CSRC, src)] for src, obj in zip(HIP_SOURCES, objs)]
        with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as ex:
            list(ex.map(lambda c: _run(c, verbose), cmds))
        tmp = out + ".tmp"
        _run([hipcc, "--offload-arch=gfx950", "-fPIC", "-shared", "-o", tmp] + objs, verbose)
        os.replace(tmp, out)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return out


def build_driver(verbose=False):
    """hmfem_run: the C host of the reservoir step (links both libraries)."""
    out = os.path.join(HERE, "hmfem_run")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    _run(["gcc", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", "-o", out, os.path.join(HERE, "driver", "hmfem_run.c"),
          "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(rocm, "include"), "-L" + HERE, "-lhmfem_gpu",
          "-lhmfem_fields", "-L" + os.path.join(rocm, "lib"), "-lamdhip64", "-lm", "-Wl,-rpath,$ORIGIN"], verbose)
    return out


def build_oracle(verbose=False):
    """The checker (oracle/liboracle.so) and, where the reference tree is
    present, the reference build oracle/_ref/hmfem_ref."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "liboracle.so"], verbose)
    if os.path.isdir("/root/reference"):
        _run(["make", "-s", "-C", odir, "ref"], verbose)


def build_all(verbose=False):
    build_fields(verbose)
    build_gpu(verbose)
    build_driver(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
