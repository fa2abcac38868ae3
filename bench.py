"""Benchmark: solves/sec and HBM GB/s of the MI355X pressure solve.

Workload (one "step" = one reservoir() pass: transmissibility assembly, AMG
setup, FP64 flexible CG with the FP32 AMG V-cycle to a relative residual of
1e-8, then flux reconstruction and divergence), by default:
  256x256x256 cells on the unit cube, K = 1e-14 plus 20 seeded karst
  conduits (random polylines of 5-15 segments, 1-3 cells thick, K x 10^U(4,6),
  seed 20261202), Dirichlet p=1 / p=0 on the x faces, no-flow elsewhere,
  lambda = 1.
--workload c2 selects the 128^3 lognormal case, c4 the 96-instance batch,
c5 the 512^3 block of eight right-hand sides, c1 the 16^3 closed-form case.
Inputs (K and the face values) are resident in HBM before timing starts.

Prints one JSON line (rank 0). Run `python bench.py --help` for options.
"""
import argparse
import json
import os
import platform
import struct
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/sec + HBM GB/s at 1e-8 rel. residual, FP32-AMG vs FP64, 1 MI355X"
HBM_PEAK_GBS = 8000.0
SEED = 20261202
WORKLOADS = {
    "c2": dict(dims=(128, 128, 128), kind="lognormal", sigma=2.0, corr=8.0,
               label="128^3 lognormal K, sigma=2, corr 8 cells, x-Dirichlet, rtol 1e-8"),
    "c1": dict(dims=(16, 16, 16), kind="homogeneous",
               label="16^3 homogeneous K=1e-14, x-Dirichlet, rtol 1e-10"),
    "c3": dict(dims=(256, 256, 256), kind="karst", conduits=20,
               label="256^3 K=1e-14 + 20 karst conduits x10^U(4,6), x-Dirichlet, rtol 1e-8"),
    "c5": dict(dims=(512, 512, 512), block=8, sigma=3.0, corr=16.0, conduits=50,
               label="512^3 lognormal K sigma=3 corr 16 cells + 50 karst conduits, all six faces Dirichlet; 8 "
                     "right-hand sides solved as one block: p=1/p=0 face pairs along x, y, z in both orientations "
                     "(linear lateral faces) and two zero-face source/sink pairs; rtol 1e-8"),
    "c4": dict(batch=96, edges=(32, 48, 64, 80, 96),
               label="96 independent instances in one batch, edges from {32,48,64,80,96} per axis; even: lognormal "
                     "sigma U[0.5,2.5] corr 8; odd: K=1e-14 + 1-5 conduits; x-Dirichlet, rtol 1e-8"),
}


def c4_instances(w):
    """Per-instance dims and K (host), drawn from the instance's own stream."""
    from hmfem_amd import fields
    out = []
    for i in range(w["batch"]):
        rs = fields.Stream(SEED, i)
        d = tuple(rs.choice(w["edges"]) for _ in range(3))
        seed = fields.instance_seed(SEED, i)
        if i % 2 == 0:
            sig = rs.uniform(0.5, 2.5)
            K = fields.lognormal(d[0], d[1], d[2], np.log(1e-14), sig, 8.0, seed)
            desc = "lognormal sigma=%.3f" % sig
        else:
            nc = 1 + min(4, int(rs.uniform() * 5))
            K = fields.conduits(fields.homogeneous(d[0], d[1], d[2], 1e-14), d[0], d[1], d[2], nc, seed)
            desc = "homogeneous + %d conduits" % nc
        out.append((d, K, desc))
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_K(w):
    from hmfem_amd import fields
    nx, ny, nz = w["dims"]
    if w["kind"] == "lognormal":
        return fields.lognormal(nx, ny, nz, np.log(1e-14), w["sigma"], w["corr"], SEED)
    K = fields.homogeneous(nx, ny, nz, 1e-14)
    if w["kind"] == "karst":
        fields.conduits(K, nx, ny, nz, w["conduits"], SEED)
    return K


def host_cpu():
    """CPU model and the cores this process may run on."""
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count()
    return model, ncpu


def write_ref_input(path, K, dims, bc_type, bc_val):
    nx, ny, nz = dims
    with open(path, "wb") as f:
        f.write(struct.pack("<3i", nx, ny, nz))
        f.write(struct.pack("<5d", 1.0, 1.0, 1.0, 1.0, 1e-12))
        f.write(struct.pack("<6i", *bc_type))
        f.write(struct.pack("<6d", *bc_val))
        f.write(struct.pack("<i", 0))
        kb = np.ascontiguousarray(K, "<f8").tobytes()
        f.write(kb); f.write(kb); f.write(kb)


def cpu_baseline_bounded(K, dims, pcg_iterations, ref_loops, budget_s, bc_type=(0, 0, 1, 1, 1, 1),
                         bc_val=(1.0, 0.0, 0.0, 0.0, 0.0, 0.0)):
    """The reference build on the full-size problem, with its PCG cut off
    after `pcg_iterations` iterations: assembly, count-CSR, AMG setup, the
    iterations and the fluxes are timed at the real size, and the PCG time is
    scaled to the `ref_loops` loops (pcg.h's _loop) the reference needs at its
    own 1e-8 stop on this problem (counted by a full run of the same build)."""
    ref = os.path.join(ROOT, "oracle", "_ref", "hmfem_ref")
    model, ncpu = host_cpu()
    nx, ny, nz = dims
    with tempfile.TemporaryDirectory() as tmp:
        inp = os.path.join(tmp, "in.bin")
        write_ref_input(inp, K, dims, bc_type, bc_val)
        t0 = time.time()
        subprocess.run([ref, inp, os.path.join(tmp, "out"), "time", str(pcg_iterations)], check=True,
                       stdout=subprocess.DEVNULL, timeout=max(120, 8 * budget_s),
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
        wall = time.time() - t0
        meta = json.load(open(os.path.join(tmp, "out.json")))
    # the PCG ran its initial residual and preconditioner application plus
    # `pcg_loop - 1` iterations; the full solve runs ref_loops - 1 iterations
    # plus the same start, so the time scales with the loop count
    t_pcg = meta["t_pcg"] * ref_loops / meta["pcg_loop"]
    t = meta["t_assemble"] + meta["t_csr"] + meta["t_setup"] + t_pcg + meta["t_flux"]
    return {"value": 1.0 / t, "unit": "solves/s", "cores": 1, "kind": "reference",
            "sample": "the reference build on the full %dx%dx%d problem, single thread: assembly %.2f s, count-CSR "
                      "%.2f s, AMG setup %.2f s, %d PCG iterations %.2f s, fluxes %.2f s (process wall %.1f s); the "
                      "PCG time is scaled to the %d loops of the reference's own 1e-8 stop on this problem (its full "
                      "run in the fixture), %.2f s, for a %.1f s step"
                      % (nx, ny, nz, meta["t_assemble"], meta["t_csr"], meta["t_setup"], meta["pcg_loop"] - 1,
                         meta["t_pcg"], meta["t_flux"], wall, ref_loops, t_pcg, t),
            "cpu": model, "host_cpus": ncpu}


def cpu_baseline(K, dims, budget_s, bc_type=(0, 0, 1, 1, 1, 1), bc_val=(1.0, 0.0, 0.0, 0.0, 0.0, 0.0)):
    """Time the reference build (oracle/_ref) on the same problem: one
    reservoir() step single-threaded. Falls back to the oracle restatement."""
    nx, ny, nz = dims
    ref = os.path.join(ROOT, "oracle", "_ref", "hmfem_ref")
    cores = 1
    model, ncpu = host_cpu()
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as tmp:
            inp = os.path.join(tmp, "in.bin")
            write_ref_input(inp, K, dims, bc_type, bc_val)
            t0 = time.time()
            subprocess.run([ref, inp, os.path.join(tmp, "out"), "time"], check=True, stdout=subprocess.DEVNULL,
                           timeout=max(60, 4 * budget_s), env=dict(os.environ, OMP_NUM_THREADS="1"))
            wall = time.time() - t0
            meta = json.load(open(os.path.join(tmp, "out.json")))
        t = meta["t_assemble"] + meta["t_solve_1e-8"] + meta["t_flux"]
        return {"value": 1.0 / t, "unit": "solves/s", "cores": cores, "kind": "reference",
                "sample": "one full %dx%dx%d step of the reference build (assembly, AMG setup + PCG to its 1e-8 "
                          "stop, fluxes), single thread; %d PCG iterations; %.2f s (process wall %.2f s)"
                          % (nx, ny, nz, meta["iter_amg_1e-8"] - 1, t, wall),
                "cpu": model, "host_cpus": ncpu}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    P = O.Problem(dims, K, bc_type, bc_val)
    t0 = time.time()
    perm, indcol, b = O.assemble(P)
    cnt, col, ele = O.build_csr(perm, indcol)
    x, it = O.compute(cnt, col, ele, b, eps=1e-8)
    O.fluxos(P, x, perm)
    t = time.time() - t0
    return {"value": 1.0 / t, "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": "one full %dx%dx%d step of the oracle restatement, single thread; %.2f s" % (nx, ny, nz, t),
            "cpu": model, "host_cpus": ncpu}


FIXTURES = {"c2": "c2_128_lognormal", "c3": "c3_256_karst"}


def reference_fixture(name, w, K):
    """The reference's full-size summary for this workload (tests/golden/full),
    if there is one for exactly this field; with the sampled cell indices."""
    if name not in FIXTURES:
        return None
    path = os.path.join(ROOT, "tests", "golden", "full", FIXTURES[name] + ".npz")
    if not os.path.exists(path):
        return None
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from summary_sample import field_digest, sample_index
    z = np.load(path)
    if tuple(int(v) for v in z["dims"]) != tuple(w["dims"]) or field_digest(K) != str(z["K_sha256"]):
        return None
    return z, sample_index(w["dims"])


def device_stream_gbs(torch, stream, mib=1024, reps=10):
    """Bytes moved per second by STREAM's Add kernel (c = a + b, FP64 arrays
    of 1 GiB each, so far past the 256 MB last-level cache) on `stream`: the
    practical HBM rate the kernels are also read against."""
    n = (mib << 20) // 8
    with torch.cuda.stream(stream):
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.ones_like(a)
        c = torch.empty_like(a)
        torch.add(a, b, out=c)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            torch.add(a, b, out=c)
        e1.record(stream)
    e1.synchronize()
    gbs = 3 * n * 8 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b, c
    return gbs


def run_batch(args, w, torch, S, emit=True):
    inst = c4_instances(w)
    dims = [d for d, _, _ in inst]
    K = np.concatenate([k for _, k, _ in inst])
    ntot = K.size
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    Kd = dev(K)
    big = max(max(a * b, b * c, a * c) for a, b, c in dims)
    bval = [dev(np.full(big, 1.0))] + [dev(np.zeros(big)) for _ in range(5)]
    p = torch.zeros(ntot, dtype=torch.float64, device="cuda")
    v = torch.zeros(ntot * 7, dtype=torch.float64, device="cuda")
    ctx = S.Context()
    solver_stream = torch.cuda.ExternalStream(ctx.stream)
    prec = S.PREC_FP64 if args.precision == "fp64" else S.PREC_FP32
    opt = S.default_options(rtol=1e-8, precision=prec, check_every=2)
    if args.smoother:
        opt.smoother = S.SMOOTH_JACOBI if args.smoother == "jacobi" else S.SMOOTH_BLOCK_GS
    for name, val in (("gs_sweeps", args.gs_sweeps), ("ocf", args.ocf), ("deflate_ratio", args.deflate_ratio)):
        if val is not None:
            setattr(opt, name, val)

    def step(o=opt):
        with torch.cuda.stream(solver_stream):
            p.zero_()
        return ctx.reservoir_batch(dims, Kd, Kd, Kd, bval, p, v, o)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st, per = res[-1]
    nb = w["batch"]
    value = nb * args.steps / elapsed
    done = np.sort(per["done_ms"])
    its = per["iterations"]
    log("batch of %d: %.3f s for %d steps, iterations %d..%d" % (nb, elapsed, args.steps, its.min(), its.max()))
    fp64 = None
    if not args.no_fp64 and prec == S.PREC_FP32:
        o64 = S.default_options(rtol=1e-8, precision=S.PREC_FP64, check_every=2)
        o64.smoother = opt.smoother
        step(o64)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        k64 = max(2, args.steps // 2)
        for _ in range(k64):
            step(o64)
        torch.cuda.synchronize()
        fp64 = {"value": nb * k64 / (time.perf_counter() - ta)}
        fp64["speedup_fp32_over_fp64"] = value / fp64["value"]
    # per-instance options from the batch autotuner (opt.autotune: tuned on
    # the first call, which is not timed)
    tuned = None
    if getattr(args, "tune", not args.no_fp64) and prec == S.PREC_FP32:
        ot = S.default_options(rtol=1e-8, check_every=2, autotune=1)
        ot.smoother = opt.smoother
        t_tune = time.perf_counter()
        step(ot)
        torch.cuda.synchronize()
        t_tune = time.perf_counter() - t_tune
        ta = time.perf_counter()
        kt = max(3, args.steps)
        for _ in range(kt):
            stt, pert = step(ot)
        torch.cuda.synchronize()
        t_tuned = time.perf_counter() - ta
        dt = np.sort(pert["done_ms"])
        gs_i, ocf_i, _ = ctx.autotune_batch(dims, Kd, Kd, Kd, bval, None, S.default_options(rtol=1e-8, check_every=2))
        pairs = {}
        for a, b in zip(gs_i.tolist(), np.round(ocf_i, 3).tolist()):
            pairs["%d sweeps, scaling %g" % (a, b)] = pairs.get("%d sweeps, scaling %g" % (a, b), 0) + 1
        tuned = {"value": nb * kt / t_tuned, "tuning_s": t_tune,
                 "per_instance_done_ms": {"p50": float(np.percentile(dt, 50)), "p99": float(np.percentile(dt, 99)),
                                          "max": float(dt[-1])},
                 "iterations": {"min": int(pert["iterations"].min()), "max": int(pert["iterations"].max())},
                 "instances_per_option": pairs}
        tuned["speedup_over_defaults"] = tuned["value"] / value
    cpu = None
    if not args.no_cpu:
        # the reference build on the batch's first instances, up to the budget
        tot, cnt, t_start = 0.0, 0, time.time()
        for d, k, _ in inst:
            r = cpu_baseline(k, d, args.cpu_budget)
            tot += 1.0 / r["value"]
            cnt += 1
            if time.time() - t_start > args.cpu_budget:
                break
        cpu = {"value": cnt / tot, "unit": "solves/s", "cores": 1, "kind": r["kind"],
               "sample": "the first %d instances of the batch, one after the other, single thread, %.1f s" % (cnt, tot)}
    out = {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64" if prec == S.PREC_FP32 else "f64",
        "data": "synthetic",
        "config": {"workload": w["label"], "instances": nb, "cells": int(ntot), "rtol": 1e-8,
                   "smoother": "block red-black Gauss-Seidel" if opt.smoother == S.SMOOTH_BLOCK_GS
                   else "damped Jacobi"},
        "per_instance_done_ms": {"p50": float(np.percentile(done, 50)), "p99": float(np.percentile(done, 99)),
                                 "max": float(done[-1])},
        "iterations": {"min": int(its.min()), "median": float(np.median(its)), "max": int(its.max())},
        "max_rel_residual": float(per["rel_residual"].max()),
        "phase_ms": {"assemble": st.t_assemble_ms, "setup": st.t_setup_ms, "solve": st.t_solve_ms,
                     "flux": st.t_flux_ms},
        "fp64_hierarchy": fp64, "autotuned": tuned, "cpu_baseline": cpu,
    }
    if emit:
        print(json.dumps(out))
    ctx.close()
    del Kd, p, v
    torch.cuda.empty_cache()
    return out


def batch_summary(torch, S):
    """The 96-instance batch workload, short (1 warmup + 3 timed batches, no
    CPU or FP64 legs; the autotuned options' line): solves/s and per-instance
    completion percentiles, for the main line."""
    a = argparse.Namespace(steps=3, warmup=1, no_fp64=True, no_cpu=True, precision="fp32", smoother=None,
                           gs_sweeps=None, ocf=None, deflate_ratio=None, cpu_budget=0.0, tune=True)
    r = run_batch(a, WORKLOADS["c4"], torch, S, emit=False)
    return {"workload": r["config"]["workload"], "value": r["value"], "unit": "solves/s",
            "ms_per_batch": r["ms_per_step"], "per_instance_done_ms": r["per_instance_done_ms"],
            "iterations": r["iterations"], "setup_ms": r["phase_ms"]["setup"], "autotuned": r["autotuned"]}


def block_rhs(nx, ny, nz, nrhs):
    """The block workload's right-hand sides: (six face arrays, b-only source or None)."""
    from hmfem_amd import fields
    out = []
    for r in range(nrhs):
        if r < 6:
            out.append((fields.dirichlet_pair_faces(nx, ny, nz, r // 2, r % 2 == 0), None))
        else:
            faces = [np.zeros(sz) for sz in fields.face_index_sizes(nx, ny, nz)]
            out.append((faces, fields.source_sink_pair(nx, ny, nz, SEED + r, 1.0 / (nx * ny * nz))))
    return out


def block_bytes_per_iteration(dims, nr):
    """Algorithmic HBM bytes of one block CG iteration on the fine level
    (coefficients once, nr interleaved right-hand sides): SpMV (A64, s, z in;
    s, v out), update (x, s, r, v in; x, r and its FP32 copy out),
    pre-smoothing (wx, wy, wz, di, FP32 r in; u out), residual-restriction
    (A32, FP32 r, u in), post-smoothing (wx, wy, wz, di, FP32 r, u, v in; z
    out). Coarse levels add about 1/7 and are left out."""
    n = dims[0] * dims[1] * dims[2]
    per = (32 + nr * (8 + 4 + 8 + 8)) + nr * 52 + (16 + nr * (4 + 4)) + (16 + nr * (4 + 4)) + \
        (16 + nr * (4 + 4 + 8 + 4))
    return n * per


def run_block(args, w, torch, S):
    from hmfem_amd import fields
    nx, ny, nz = w["dims"]
    n = nx * ny * nz
    nr = w["block"]
    log("workload c5", w["dims"], "generating K")
    K = fields.lognormal(nx, ny, nz, np.log(1e-14), w["sigma"], w["corr"], SEED)
    fields.conduits(K, nx, ny, nz, w["conduits"], SEED + 1)
    log("K ready; right-hand sides")
    rhs = block_rhs(nx, ny, nz, nr)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    Kd = dev(K)
    bvals = [[dev(a) for a in faces] for faces, _ in rhs]
    qs = [None if q is None else dev(q) for _, q in rhs]
    ps = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(nr)]
    vs = [torch.zeros(n * 7, dtype=torch.float64, device="cuda") for _ in range(nr)]
    g = S.Context.grid(nx, ny, nz, None, 1.0, (0, 0, 0, 0, 0, 0))
    ctx = S.Context()
    solver_stream = torch.cuda.ExternalStream(ctx.stream)
    opt = S.default_options(rtol=1e-8)
    for name, val in (("gs_sweeps", args.gs_sweeps), ("ocf", args.ocf), ("deflate_ratio", args.deflate_ratio)):
        if val is not None:
            setattr(opt, name, val)
    if args.no_compact:
        opt.block_compact = 0

    def step():
        with torch.cuda.stream(solver_stream):
            for p in ps:
                p.zero_()
        return ctx.reservoir_block(g, Kd, Kd, Kd, bvals, ps, vs, qs, opt)

    for q in range(args.warmup):
        st, per = step()
        log("warmup block step %d: iterations %s, %.0f ms" % (q, per["iterations"].tolist(), st.t_solve_ms))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = []
    for q in range(args.steps):
        res.append(step())
        log("timed block step %d done" % q)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st, per = res[-1]
    value = nr * args.steps / elapsed
    its = per["iterations"]
    # the library runs the right-hand sides through four lanes of one interleaved solve
    gbs = block_bytes_per_iteration(w["dims"], 4) * st.block_iterations / (st.t_solve_ms * 1e-3) / 1e9
    log("block: %.3f s per step, iterations %s" % (elapsed / args.steps, its.tolist()))

    # the same eight systems through the single-right-hand-side path, one after the other
    separate = None
    if not args.no_fp64:
        o1 = S.default_options(rtol=1e-8)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        perm = indcol = b = None
        torch.cuda.synchronize()
        ta = time.perf_counter()
        its1 = []
        for r, (bv, q) in enumerate(zip(bvals, qs)):
            with torch.cuda.stream(solver_stream):
                x.zero_()
            if q is None:
                s1 = ctx.reservoir(g, Kd, Kd, Kd, bv, x, vs[r], o1)
            else:
                if perm is None:
                    perm = torch.zeros(n * 7, dtype=torch.float64, device="cuda")
                    indcol = torch.zeros(n * 7, dtype=torch.int32, device="cuda")
                    b = torch.zeros(n, dtype=torch.float64, device="cuda")
                ctx.assemble(g, Kd, Kd, Kd, bv, b, perm, indcol)
                b += q
                s1 = ctx.solve_structured(nx, ny, nz, perm, indcol, b, x, o1)
                ctx.flux(g, Kd, Kd, Kd, bv, x, vs[r])
            its1.append(s1.iterations)
        torch.cuda.synchronize()
        t1 = time.perf_counter() - ta
        separate = {"value": nr / t1, "unit": "solves/s", "iterations": its1,
                    "speedup_block_over_separate": value / (nr / t1)}
        log("separate: %.3f s" % t1)

    cpu = None
    if not args.no_cpu:
        # the reference on a 128^3 member of the same family (sigma 3, correlation
        # scaled to 4 cells, 50/64 conduits rounded up), extrapolated per unknown
        m = 128
        Ks = fields.lognormal(m, m, m, np.log(1e-14), w["sigma"], w["corr"] * m / nx, SEED)
        fields.conduits(Ks, m, m, m, max(1, (w["conduits"] * m * m) // (nx * nx)), SEED + 1)
        r = cpu_baseline(Ks, (m, m, m), args.cpu_budget, (0, 0, 0, 0, 0, 0), (1.0, 0.0, 0.0, 0.0, 0.0, 0.0))
        scale = (n / m ** 3) * nr
        cpu = {"value": nr * r["value"] / scale, "unit": "solves/s", "cores": 1, "kind": r["kind"],
               "sample": "extrapolated: " + r["sample"] + "; scaled by %d (cells) per right-hand side, iteration "
                         "growth with size not included" % (n // m ** 3)}
    out = {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64", "data": "synthetic",
        "config": {"workload": w["label"], "cells": n, "right_hand_sides": nr, "rtol": 1e-8,
                   "precision": "FP32 AMG in FP64 flexible CG, %d right-hand sides through 4 interleaved lanes%s"
                   % (nr, "" if args.no_compact else ", 2 once no right-hand side is left to start")},
        "iterations": its.tolist(), "block_iterations": st.block_iterations,
        "max_rel_residual": float(per["rel_residual"].max()),
        "div_max": per["div_max"].tolist(),
        "phase_ms": {"assemble": st.t_assemble_ms, "setup": st.t_setup_ms, "solve": st.t_solve_ms,
                     "flux": st.t_flux_ms},
        "hbm_gbps_solve_phase_fine_level": gbs,
        "footprint_gb": st.footprint_bytes / 1e9, "hbm_capacity_gb": 288.0,
        "separate_solves": separate, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 20; c3: 10; c4: 3; c5: 1)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps first (default 3; c3: 2; c4, c5: 1)")
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp16"],
                    help="hierarchy precision; fp16 = FP32 hierarchy with the fine level stored in FP16")
    ap.add_argument("--smoother", default=None, choices=["jacobi", "blockgs"],
                    help="smoother of the multi-block levels (default: the library's)")
    ap.add_argument("--gs-sweeps", type=int, default=None, help="block Gauss-Seidel sweeps (default: the library's)")
    ap.add_argument("--ocf", type=float, default=None, help="coarse-correction scaling (default: the library's)")
    ap.add_argument("--deflate-ratio", type=float, default=None,
                    help="conduit correction: strong-cell threshold over the median diagonal (default: the library's)")
    ap.add_argument("--no-compact", action="store_true",
                    help="c5: keep four lanes to the end (no switch to two once the feed is empty)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison")
    ap.add_argument("--profile-only", action="store_true", help="timed steps only (for rocprofv3 runs)")
    ap.add_argument("--no-batch", action="store_true",
                    help="c3: leave out the short run of the batch workload (solves/s, per-instance p50/p99)")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--edge", type=int, default=0, metavar="N",
                    help="smaller cube edge for the c2/c3/c5 workload (profiling runs; conduit counts scale with "
                         "the cross-section); the JSON names the size")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (counter calibration)")
    args = ap.parse_args()
    big = {"c3": (10, 2), "c4": (3, 1), "c5": (1, 1)}.get(args.workload, (20, 3))
    if args.steps is None:
        args.steps = big[0]
    if args.warmup is None:
        args.warmup = big[1]
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every box here has one); --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from hmfem_amd import solver as S

    if not torch.cuda.is_available():
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.calibrate_copy:
        nb = args.calibrate_copy << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": nb, "launches": 5}))
        return
    w = WORKLOADS[args.workload]
    if args.edge and "dims" in w:
        e = args.edge
        nx0 = w["dims"][0]
        w = dict(w, dims=(e, e, e), label="%s [cube edge %d instead of %d]" % (w["label"], e, nx0))
        if "conduits" in w:
            w["conduits"] = max(1, round(w["conduits"] * e * e / (nx0 * nx0)))
    if "batch" in w:
        return run_batch(args, w, torch, S)
    if "block" in w:
        return run_block(args, w, torch, S)
    nx, ny, nz = w["dims"]
    n = nx * ny * nz
    rtol = 1e-10 if args.workload == "c1" else 1e-8
    log("workload", args.workload, w["dims"], "generating K")
    K = make_K(w)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    Kd = dev(K)
    faces = [np.full(ny * nz, 1.0), np.zeros(ny * nz), np.zeros(nx * nz), np.zeros(nx * nz), np.zeros(nx * ny),
             np.zeros(nx * ny)]
    bval = [dev(a) for a in faces]
    p = torch.zeros(n, dtype=torch.float64, device="cuda")
    v = torch.zeros(n * 7, dtype=torch.float64, device="cuda")
    g = S.Context.grid(nx, ny, nz, None, 1.0, (0, 0, 1, 1, 1, 1))
    ctx = S.Context()
    prec = {"fp64": S.PREC_FP64, "fp16": S.PREC_FP16}.get(args.precision, S.PREC_FP32)
    opt = S.default_options(rtol=rtol, precision=prec)
    if args.smoother:
        opt.smoother = S.SMOOTH_JACOBI if args.smoother == "jacobi" else S.SMOOTH_BLOCK_GS
    for name, val in (("gs_sweeps", args.gs_sweeps), ("ocf", args.ocf), ("deflate_ratio", args.deflate_ratio)):
        if val is not None:
            setattr(opt, name, val)

    solver_stream = torch.cuda.ExternalStream(ctx.stream)

    def step(o=opt):
        with torch.cuda.stream(solver_stream):
            p.zero_()   # each step solves from the reference's zero initial guess
        return ctx.reservoir(g, Kd, Kd, Kd, bval, p, v, o)

    for _ in range(args.warmup):
        st = step()
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(step())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = args.steps / elapsed
    st = stats[-1]
    log("timed", args.steps, "steps in %.3f s" % elapsed, "iterations", st.iterations, "levels", st.levels)
    if args.profile_only:
        print(json.dumps({"metric": METRIC, "value": value, "unit": "solves/s", "steps": args.steps}))
        return

    # Per-kernel durations with HIP events on the solver's own stream.
    ctx.profile_enable(40000)
    prof_steps = max(2, min(5, args.steps))
    for _ in range(prof_steps):
        step()
    prof = ctx.profile_read()
    ctx.profile_enable(0)
    fine = {k: d for k, d in prof.items() if d["count"] > 0 and k != "coarse"}
    dom = max(fine, key=lambda k: fine[k]["ms"])
    d = fine[dom]
    avg_ms = d["ms"] / d["count"]
    achieved = d["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9
    kernels = {k: {"avg_us": 1e3 * x["ms"] / x["count"], "launches_per_step": x["count"] / prof_steps,
                   "bytes_per_launch": x["bytes_per_launch"] or None,
                   "gbps": x["bytes_per_launch"] / (x["ms"] / x["count"] * 1e-3) / 1e9 if x["bytes_per_launch"] else None}
               for k, x in prof.items() if x["count"] > 0}
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload)
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get(dom)
    stream_gbs = device_stream_gbs(torch, solver_stream)

    # Plain FP64 hierarchy on the same workload (FP32-AMG against FP64).
    fp64 = None
    if not args.no_fp64 and prec == S.PREC_FP32:
        o64 = S.default_options(rtol=rtol, precision=S.PREC_FP64)
        o64.smoother = opt.smoother
        step(o64)
        torch.cuda.synchronize()
        k64 = max(3, args.steps // 2)
        ta = time.perf_counter()
        for _ in range(k64):
            s64 = step(o64)
        torch.cuda.synchronize()
        fp64 = {"value": k64 / (time.perf_counter() - ta), "iterations": s64.iterations}
        fp64["speedup_fp32_over_fp64"] = value / fp64["value"]
        # accuracy of the FP32-AMG solution against an FP64 solve driven to 1e-12
        # the copies run on the solver's stream: the next step's p.zero_() is queued there too
        step(opt)
        with torch.cuda.stream(solver_stream):
            p32 = p.clone()
        step(S.default_options(rtol=1e-12, precision=S.PREC_FP64))
        with torch.cuda.stream(solver_stream):
            pf = p.clone()
            fp64["rel_l2_fp32amg_vs_fp64_1e-12"] = float(torch.linalg.norm(p32 - pf) / torch.linalg.norm(pf))

    # Plain FP64 PCG (diagonal preconditioner, no hierarchy) on the same workload
    plain = None
    if not args.no_fp64 and prec == S.PREC_FP32:
        oj = S.default_options(rtol=rtol, precision=S.PREC_FP64)
        oj.precond = S.PRECOND_JACOBI
        oj.max_iterations = 50000
        sj = step(oj)
        torch.cuda.synchronize()
        kj = 2
        ta = time.perf_counter()
        for _ in range(kj):
            sj = step(oj)
        torch.cuda.synchronize()
        plain = {"value": kj / (time.perf_counter() - ta), "iterations": sj.iterations,
                 "converged": sj.status == S.OK, "rel_residual": sj.rel_residual}
        plain["speedup_fp32amg_over_plain"] = value / plain["value"]

    # repeated solves that reuse the hierarchy and conduit correction of the
    # previous call (opt.reuse_setup; here K does not change between calls)
    reuse = None
    if not args.no_fp64 and prec == S.PREC_FP32:
        orr = S.default_options(rtol=rtol, precision=prec)
        orr.smoother = opt.smoother
        orr.reuse_setup = 1
        step(orr)
        torch.cuda.synchronize()
        kr = max(3, args.steps // 2)
        ta = time.perf_counter()
        for _ in range(kr):
            sr = step(orr)
        torch.cuda.synchronize()
        reuse = {"value": kr / (time.perf_counter() - ta), "iterations": sr.iterations, "setup_ms": sr.t_setup_ms}
        reuse["speedup_over_rebuild"] = reuse["value"] / value

    # the autotuner's choice for this problem (tuning time is not in any figure)
    tuned = None
    if not args.no_fp64 and prec == S.PREC_FP32:
        ot, tms = ctx.autotune(g, Kd, Kd, Kd, bval, S.default_options(rtol=rtol, precision=prec))
        step(ot)
        torch.cuda.synchronize()
        kt = max(3, args.steps // 2)
        ta = time.perf_counter()
        for _ in range(kt):
            stt = step(ot)
        torch.cuda.synchronize()
        tuned = {"value": kt / (time.perf_counter() - ta), "iterations": stt.iterations,
                 "options": {"smoother": "block GS" if ot.smoother == S.SMOOTH_BLOCK_GS else "Jacobi",
                             "gs_sweeps": ot.gs_sweeps, "nu": ot.nu, "ocf": ot.ocf, "tail_gamma": ot.tail_gamma}}

    # FP16-stored fine level (ablation) on the same workload
    fp16 = None
    if not args.no_fp64 and prec == S.PREC_FP32 and opt.smoother == S.SMOOTH_BLOCK_GS:
        o16 = S.default_options(rtol=rtol, precision=S.PREC_FP16)
        step(o16)
        torch.cuda.synchronize()
        k16 = max(3, args.steps // 2)
        ta = time.perf_counter()
        for _ in range(k16):
            s16 = step(o16)
        torch.cuda.synchronize()
        fp16 = {"value": k16 / (time.perf_counter() - ta), "iterations": s16.iterations,
                "rel_residual": s16.rel_residual}
        fp16["speedup_over_fp32"] = fp16["value"] / value

    # accuracy against the reference's own pressure on this problem (its 1e-12
    # solve, sampled; tests/golden/full), for the timed 1e-8 solve and at 1e-9
    accuracy = None
    fx = reference_fixture(args.workload, w, K)
    if fx is not None:
        z, idx = fx
        ref_s = z["p12_s"]
        rel = lambda x: float(np.linalg.norm(x[idx] - ref_s) / np.linalg.norm(ref_s))
        step(opt)
        e8 = rel(p.cpu().numpy())
        step(S.default_options(rtol=rtol * 0.1, precision=prec, smoother=opt.smoother))
        e9 = rel(p.cpu().numpy())
        step(S.default_options(rtol=1e-12, precision=prec, smoother=opt.smoother))
        e12 = rel(p.cpu().numpy())
        accuracy = {"rel_l2_vs_reference_1e-12": e8, "rtol": rtol,
                    "reference_own_1e-8_vs_its_1e-12": float(z["rel_l2_p8_p12"]),
                    "rel_l2_vs_reference_at_rtol_%g" % (rtol * 0.1): e9,
                    "rel_l2_vs_reference_at_rtol_1e-12": e12,
                    "reference_iterations_1e-8": int(z["iter8"]) - 1,
                    "samples": int(idx.size),
                    "note": "relative L2 over the fixture's cells (every 8th cell per axis and the three mid-planes) "
                            "of the GPU pressure against the reference build's pressure at 1e-12"}
        log("accuracy", accuracy)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        m = 128
        if n <= m ** 3:
            cpu = cpu_baseline(K, w["dims"], args.cpu_budget)
        elif fx is not None:
            cpu = cpu_baseline_bounded(K, w["dims"], 5, int(fx[0]["iter8"]), args.cpu_budget)
        else:
            # a smaller cube of the same family with no reference run at this
            # size: the 128^3 member, scaled per unknown and labelled so
            ws = dict(w, dims=(m, m, m))
            if "conduits" in ws:
                ws["conduits"] = max(1, round(w["conduits"] * m * m / (nx * ny)))
            r = cpu_baseline(make_K(ws), (m, m, m), args.cpu_budget)
            f = n / m ** 3
            cpu = dict(r, value=r["value"] / f,
                       sample="extrapolated: " + r["sample"] + "; %d conduits; time scaled by %g (cells), iteration "
                              "growth with size not included" % (ws.get("conduits", 0), f))
        cpu["cores"] = 1

    solve_gbs = st.solve_bytes / (st.t_solve_ms * 1e-3) / 1e9 if st.t_solve_ms > 0 else None
    out = {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": {S.PREC_FP32: "f32/f64", S.PREC_FP16: "f16/f32/f64"}.get(prec, "f64"),
        "data": "synthetic",
        "config": {"workload": w["label"], "cells": n, "precision": {
                       S.PREC_FP32: "FP32 AMG in FP64 flexible CG",
                       S.PREC_FP16: "FP32 AMG with an FP16-stored fine level in FP64 flexible CG"}.get(
                       prec, "FP64 AMG in FP64 flexible CG"), "rtol": rtol,
                   "smoother": "block red-black Gauss-Seidel" if opt.smoother == S.SMOOTH_BLOCK_GS
                   else "damped Jacobi"},
        "iterations": st.iterations, "levels": st.levels,
        "deflation": {"clusters": st.deflation_clusters, "cells": st.deflation_cells},
        "phase_ms": {"assemble": st.t_assemble_ms, "setup": st.t_setup_ms, "solve": st.t_solve_ms,
                     "flux": st.t_flux_ms},
        "hbm_gbps_solve_phase": solve_gbs,
        "rel_residual": st.rel_residual, "div_max": st.div_max,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "stream_add_gbs": stream_gbs, "frac_of_stream_add": achieved / stream_gbs,
                     "bytes_per_launch": d["bytes_per_launch"], "avg_launch_us": 1e3 * avg_ms},
        "kernels": kernels,
        "accuracy_vs_reference": accuracy,
        "fp64_hierarchy": fp64,
        "plain_fp64_jacobi_pcg": plain,
        "reuse_setup": reuse,
        "fp16_fine_level": fp16,
        "autotuned": tuned,
        "cpu_baseline": cpu,
    }
    if args.workload == "c3" and not args.no_batch:
        log("batch workload (c4), short")
        out["batch"] = batch_summary(torch, S)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
