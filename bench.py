#!/usr/bin/env python3
"""FDFD serving benchmark on one MI355X.

Workload (the serving stream): Ez requests on square grids drawn from
{128, 192, 256, 384, 512, 768} with NPML 15-25 scaled to the grid, lambda
~ U[1.3, 1.7] um, dl = 0.02, a seeded Bernoulli design region and 1-4 point
sources each, arriving as a Poisson process. A step is one batch of
--batch requests of that stream; --steps batches make the timed stream
(8 x 64 = the 512-request stream by default).

Metric: solves/s (right-hand sides solved to relative residual <= 1e-10,
each the equivalent of one reference solve_fields call) with p50/p99
latency, for the complex64-factor + FP64-refinement path (the headline
value) and for the FP64-factor path on the same trace.

Two loads: the throughput leg offers the stream faster than it is served
(Poisson arrivals at --rate), so solves/s is the saturated rate and its
latencies include queueing; the latency leg replays the first
--latency-requests requests at --latency-load (0.25) x the FP64 path's saturated request rate,
below both paths' capacity at that load (the factorizations batch less when
fewer requests are in flight, so capacity falls with the load), and its
p50/p99 are the headline latencies.

Also reported: the roofline of the dominant kernel (the Gauss-Jordan step
of the complex64 factor, 8 factorizations per launch as in the stream,
timed with HIP events on its stream; HBM bytes per launch from two
rocprofv3 --pmc passes over the same launches, scaled by a copy of known
size) and a CPU baseline: the oracle's SuperLU restatement of the
reference's 'scipy' branch in a pool of one worker per host core, replaying
a bounded prefix of the same stream.
"""

import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_MFMA_PEAK_TFLOPS = 157.3
FP64_MFMA_PEAK_TFLOPS = 78.6
ROOF_N = 512          # grid of the roofline leg's factorizations
ROOF_JOBS = 10        # factorizations per step launch: 20 chains (the stream's launches carry 18-22)
COPY_BYTES = 256 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed batches of the stream")
    ap.add_argument("--warmup", type=int, default=1, help="untimed batches (of 8 requests each)")
    ap.add_argument("--batch", type=int, default=64, help="requests per step")
    ap.add_argument("--rate", type=float, default=200.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--slots", type=int, default=48, help="requests in flight")
    ap.add_argument("--engines", type=int, default=2, help="factorization engines (streams of step launches)")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--cpu-seconds", type=float, default=25.0,
                    help="wall-time budget of the CPU pool's sample of the stream")
    ap.add_argument("--cpu-workers", type=int, default=0,
                    help="CPU pool size (0: the host cores of this process, at most 16)")
    ap.add_argument("--latency-requests", type=int, default=160,
                    help="requests of the latency leg (arrivals below capacity)")
    ap.add_argument("--latency-load", type=float, default=0.25,
                    help="arrival rate of the latency leg as a fraction of the FP64 path's request rate")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 byte-counter passes")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--stream-only", action="store_true",
                    help="only the throughput leg of the stream (for kernel traces of it)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the kernels of the roofline leg (both precisions) and a 512 x 512 solve with "
                         "4 right-hand sides in both precisions (for counter runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--roofline-only", action="store_true", help="only the dominant-kernel measurement")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only a plain 256 MiB device copy (scale for the PMC byte counters)")
    ap.add_argument("--configs", action="store_true",
                    help="only the per-workload timings of configs 1, 2, 4 and 5 (one JSON object)")
    return ap.parse_args()


def _timed_factor(op, precision, stepwise):
    import torch
    op.factor(precision, stepwise=stepwise)          # warm
    op.drop_factor()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(s)
    op.factor(precision, stepwise=stepwise)
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def _timed_wrap(op, precision):
    """The x-wrap correction step alone (fdfd_factor_wrap) on op's resident
    factor, timed with HIP events on its stream."""
    import ctypes
    import torch
    from fdfdpy_amd import _lib as L
    from fdfdpy_amd.device import _stream, _vp
    lib = L.lib()
    g = ctypes.byref(op.grid)
    fac = op.factors[precision][0]
    work = torch.empty(lib.fdfd_factor_workspace_bytes(g, L.PREC[precision]), dtype=torch.uint8, device="cuda:0")
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(s)
    L.check(lib.fdfd_factor_wrap(g, _vp(op.op), _vp(fac), L.PREC[precision], _vp(work), _stream("cuda:0")))
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def _roof_ops(n, njobs):
    """njobs operators of the stream's n x n grid (request seeds 7, 8, ...)."""
    from fdfdpy_amd import configs as C
    from fdfdpy_amd.device import DeviceOperator
    ops = []
    for j in range(njobs):
        p = C.c3_request(7 + j, n=n)
        op = DeviceOperator(p.shape, p.omega, p.dl, p.npml, p.pol, p.L0, "cuda:0")
        op.set_eps(p.eps)
        ops.append(op)
    return ops


def _batched_factor(ops, precision, stream_timed=True):
    """fdfd_factor_batch over ops (one grid) on the current stream, twice:
    warm, then timed with HIP events on that stream. Returns ms."""
    import ctypes
    import torch
    from fdfdpy_amd import _lib as L
    from fdfdpy_amd.device import _stream
    lib = L.lib()
    g = ctypes.byref(ops[0].grid)
    pr = L.PREC[precision]
    nj = len(ops)
    facs = [torch.empty(lib.fdfd_factor_bytes(g, pr), dtype=torch.uint8, device="cuda:0") for _ in ops]
    works = [torch.empty(lib.fdfd_factor_workspace_bytes(g, pr), dtype=torch.uint8, device="cuda:0") for _ in ops]
    arr = lambda ts: (ctypes.c_void_p * nj)(*[t.data_ptr() for t in ts])
    a_ops, a_fac, a_work = arr([o.op for o in ops]), arr(facs), arr(works)
    ms = None
    for rep in range(2 if stream_timed else 1):
        s = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(s)
        L.check(lib.fdfd_factor_batch(g, a_ops, a_fac, nj, pr, a_work, _stream("cuda:0")))
        e1.record(s)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
    for w in works:
        L.check(lib.fdfd_factor_status(g, pr, ctypes.c_void_p(w.data_ptr()), _stream("cuda:0")))
    for op, f in zip(ops, facs):
        op.factors[precision] = (f, op.version, False)
    return ms


def roofline_leg(n=ROOF_N, precision="c64", njobs=ROOF_JOBS, traffic=None):
    """The dominant kernel of the stream, k_gj_step, at the stream's batching.

    njobs twisted factorizations of n x n operators share each step launch
    (fdfd_factor_batch: every launch carries a step of each operator's top
    chain and, until it ends, of its bottom chain; (c + 1) n/32 + 1 launches
    for twist block c), timed with HIP events on their stream; the x-wrap
    corrections that follow them are timed alone and taken out. Algorithmic
    flops: 8 m n^3 per factorization (m = n blocks), spread over the
    launches. One factorization alone (its two chains per launch) is timed
    too, and, for comparison, the one-chain factorization (stepwise and the
    persistent single-launch kernel), both with their wrap corrections.
    """
    import ctypes
    from fdfdpy_amd import _lib as L
    lib = L.lib()
    ops = _roof_ops(n, njobs)
    c = lib.fdfd_twist_block(ctypes.byref(ops[0].grid))
    ms_all = _batched_factor(ops, precision)
    ms_wrap = sum(_timed_wrap(op, precision) for op in ops)
    ms = ms_all - ms_wrap
    one = ops[0]
    ms1 = _timed_factor(one, precision, stepwise=True) - ms_wrap / njobs
    ms_single = _timed_factor(one, precision, stepwise=False)          # with its wrap correction
    prev = lib.fdfd_set_twist(0)
    try:
        ms_chain = _timed_factor(one, precision, stepwise=True)        # one chain, with its wrap correction
        ms_persist = _timed_factor(one, precision, stepwise=False)
        one.drop_factor()
    finally:
        lib.fdfd_set_twist(prev)
    nt = (n + 31) // 32
    launches = (c + 1) * nt + 1
    flops = 8.0 * n ** 4                       # 8 m n^3 real flops per factorization, m = n blocks
    avg_us = ms * 1e3 / launches
    achieved = njobs * flops / (ms * 1e-3) / 1e12
    peak = FP32_MFMA_PEAK_TFLOPS if precision == "c64" else FP64_MFMA_PEAK_TFLOPS
    tname = "float" if precision == "c64" else "double"
    out = {
        "kernel": f"k_gj_step<{tname}, false>",
        "bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
        "frac": round(achieved / peak, 5), "traffic": None,
        "workload": f"{njobs} twisted complex factors of {n}x{n} Ez operators (twist block {c}) sharing "
                    f"{launches} step launches: {njobs * flops:.3e} flops, {njobs * flops / launches:.3e} per launch",
        "avg_launch_us": round(avg_us, 3), "batched_factor_ms": round(ms, 3), "wrap_ms": round(ms_wrap, 3),
        "hbm_gbs_algorithmic": None,
        "one_request": {"avg_launch_us": round(ms1 * 1e3 / launches, 3), "factor_ms": round(ms1, 3),
                        "achieved_tflops": round(flops / (ms1 * 1e-3) / 1e12, 3),
                        "factor_with_wrap_ms": round(ms_single, 3)},
        "one_chain": {"stepwise_with_wrap_ms": round(ms_chain, 3), "persistent_with_wrap_ms": round(ms_persist, 3)},
    }
    # algorithmic bytes: per block step, the upper tiles of the working matrix
    # read and written once (symmetric steps), plus the tiles below the
    # diagonal written at each block's last step; n blocks of nt steps per
    # factorization, over the launches
    es = 8 if precision == "c64" else 16
    upper = nt * (nt + 1) // 2 * 32 * 32 * es * 2
    lower = nt * (nt - 1) // 2 * 32 * 32 * es / nt
    alg = njobs * n * nt * (upper + lower) / launches
    out["algorithmic_bytes_per_launch"] = round(alg)
    out["hbm_gbs_algorithmic"] = round(alg / (avg_us * 1e-6) / 1e9, 1)
    if traffic:
        out["traffic"] = traffic.get("bytes_per_launch")
        out["traffic_detail"] = traffic
    return out


def traffic_child(n=ROOF_N, precision="c64", njobs=ROOF_JOBS):
    """Run under rocprofv3 --pmc: a plain copy of known size (three times),
    then the roofline leg's batched factorization once."""
    import torch
    a = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    _batched_factor(_roof_ops(n, njobs), precision, stream_timed=False)
    print("traffic child done", flush=True)


def kernels_only(n=ROOF_N, njobs=ROOF_JOBS, nrhs=4):
    """The copy calibration, the roofline leg's batched factorization in both
    precisions, and one n x n solve with nrhs right-hand sides to 1e-10 in
    both precisions (persistent sweeps, wrap correction, FP64 stencil
    residuals): the kernels whose counters the profiles report."""
    import torch
    copy_calibration()
    ops = _roof_ops(n, njobs)
    rng = torch.Generator(device="cpu").manual_seed(1)
    out = {}
    for prec in ("c64", "fp64"):
        out[prec + "_batched_factor_ms"] = round(_batched_factor(ops, prec, stream_timed=False), 3)
        b = torch.randn((nrhs,) + ops[0].shape, dtype=torch.complex128, generator=rng).to("cuda:0")
        x, rep = ops[0].solve(b, precision=prec, tol=1e-10)
        out[prec + "_solve"] = {"steps": rep.iterations, "sweeps": rep.sweeps, "residual": rep.max_rel_residual}
        for op in ops:
            op.drop_factor()
    torch.cuda.synchronize()
    return out


def _pmc_means(counter, timeout_s=240):
    """Mean per-dispatch value of one counter (KiB) for the step kernel and
    the copy kernel, from one rocprofv3 --pmc pass over traffic_child()."""
    rp = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(rp):
        raise RuntimeError("rocprofv3 not found")
    d = tempfile.mkdtemp(prefix="fdfd_pmc_")
    try:
        cmd = [rp, "--pmc", counter, "-d", d, "-o", "pmc", "--output-format", "csv", "--",
               sys.executable, os.path.abspath(__file__), "--traffic-child"]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=timeout_s,
                           env=dict(os.environ, TMPDIR=d))
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            tail = r.stdout.decode(errors="replace")[-400:]
            raise RuntimeError(f"rocprofv3 wrote no counter file (rc {r.returncode}): {tail}")
        per = {}
        for f in files:
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") != counter:
                        continue
                    key = (row["Kernel_Name"], row.get("Dispatch_Id", ""))
                    per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
        # demangled ("k_gj_step<float, false>") or mangled ("k_gj_stepIfLb0E") names
        step = [v for (k, _), v in per.items() if "k_gj_step<float, false>" in k or "k_gj_stepIfLb0E" in k]
        copy = [v for (k, _), v in per.items() if "copyBuffer" in k]
        if not step or not copy:
            names = sorted({k for k, _ in per})[:6]
            raise RuntimeError(f"counter rows missing (step {len(step)}, copy {len(copy)}; kernels {names})")
        return sum(step) / len(step), len(step), max(copy)
    finally:
        shutil.rmtree(d, ignore_errors=True)


def traffic_leg():
    """HBM bytes per batched step launch: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) in two rocprofv3 --pmc passes, each scaled by what the counter
    reads for the 256 MiB copy. Runs before this process touches the GPU
    (the passes are child processes)."""
    try:
        f_kib, nd, f_copy = _pmc_means("FETCH_SIZE")
        w_kib, _, w_copy = _pmc_means("WRITE_SIZE")
    except Exception as e:   # reported, the rest of the bench goes on
        return {"bytes_per_launch": None, "error": f"{type(e).__name__}: {e}"[:300]}
    copy_kib = COPY_BYTES / 1024
    cf, cw = copy_kib / f_copy, copy_kib / w_copy
    return {"bytes_per_launch": round((f_kib * cf + w_kib * cw) * 1024), "dispatches": nd,
            "fetch_kib": round(f_kib, 1), "write_kib": round(w_kib, 1),
            "fetch_scale": round(cf, 3), "write_scale": round(cw, 3),
            "calibration": f"{COPY_BYTES >> 20} MiB device copy: FETCH_SIZE {f_copy:.0f} KiB, "
                           f"WRITE_SIZE {w_copy:.0f} KiB"}


def copy_calibration(nbytes=COPY_BYTES, reps=3):
    """A plain device copy of known size (scales the PMC byte counters)."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"copy_bytes_each_way": nbytes, "copies": reps}


def _cpu_solve(task):
    """One reference-style solve in a pool worker: CSR assembly + SuperLU
    (linalg.py:41-107, :140-141) for one right-hand side of a stream request."""
    seed, q = task
    from threadpoolctl import threadpool_limits
    from fdfdpy_amd import configs as C
    from oracle import construct_A, solver_direct
    with threadpool_limits(limits=1):
        p = C.c3_request(seed)
        t = time.perf_counter()
        A, _ = construct_A(p.omega, [0, p.shape[0] * p.dl], [0, p.shape[1] * p.dl], p.eps, p.npml, p.pol, p.L0)
        x = solver_direct(A, p.srcs[q] * 1j * p.omega)
        t1 = time.perf_counter()
        return seed, q, t1 - t, t1, x


def cpu_baseline(reqs, seconds, workers=0):
    """The oracle's restatement of the reference 'scipy' path (CSR assembly +
    SuperLU spsolve, refactored per right-hand side as linalg.py:122-149
    does) in a pool of one single-threaded worker per host core, replaying a
    prefix of the stream in arrival order (arrival times as in the stream).
    The prefix is the longest whose estimated work fits `seconds` of wall
    time on the pool, from one timed 192 x 192 solve and a cubic growth in n.
    Reports solves/s, p50/p99 latency (arrival -> last RHS done, queueing
    included) and p50/p99 service time (first RHS started -> last done)."""
    import multiprocessing as mp
    import numpy as np
    W = workers or min(len(os.sched_getaffinity(0)), 16)
    ctx = mp.get_context("spawn")
    with ctx.Pool(W) as pool:
        # calibration (also warms every worker's imports)
        probe = [r for r in reqs if r.problem.shape[0] == 192][:1]
        pool.map(_cpu_solve, [(0, 0)] * W)
        t192 = pool.apply(_cpu_solve, ((probe[0].rid if probe else 0, 0),))[2] if probe else 0.4
        budget = W * seconds
        sample, work = [], 0.0
        for r in reqs:
            n = r.problem.shape[0]
            w = r.problem.srcs.shape[0] * t192 * (n / 192.0) ** 3
            if sample and work + w > budget:
                break
            sample.append(r)
            work += w
        t0 = time.perf_counter()
        asyncs = []
        for r in sample:
            wait = r.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            for q in range(r.problem.srcs.shape[0]):
                asyncs.append((r, pool.apply_async(_cpu_solve, ((r.rid, q),))))
        done, sols = {}, {}
        for r, a in asyncs:
            seed, q, dt, tend, x = a.get()
            done.setdefault(r.rid, []).append((dt, tend))
            sols[(r.rid, q)] = x
        # to the last solve's end (the solutions' transfer back is not timed)
        wall = max(te for v in done.values() for _, te in v) - t0
    lat, svc = [], []
    for r in sample:
        ends = [te for _, te in done[r.rid]]
        lat.append(max(ends) - t0 - r.arrival)
        svc.append(max(ends) - min(te - dt for dt, te in done[r.rid]))
    nsolves = sum(r.problem.srcs.shape[0] for r in sample)
    sizes = sorted({r.problem.shape[0] for r in sample})
    return {"value": round(nsolves / wall, 4), "unit": "solves/s", "cores": W, "kind": "port",
            "sample": f"first {len(sample)} requests of the stream ({nsolves} solves, grids {sizes}) replayed "
                      f"at their arrival times on a pool of {W} single-threaded workers, one SuperLU refactor "
                      f"per RHS; {wall:.1f} s wall",
            "p50_latency_ms": round(float(np.percentile(lat, 50)) * 1e3, 1),
            "p99_latency_ms": round(float(np.percentile(lat, 99)) * 1e3, 1),
            "p50_service_ms": round(float(np.percentile(svc, 50)) * 1e3, 1),
            "p99_service_ms": round(float(np.percentile(svc, 99)) * 1e3, 1)}, sols


def result_errors(res, ref):
    """Relative L2 error of each result (request, RHS) of `res` against
    `ref`: {(rid, q): solution} (device tensors or flattened numpy arrays of
    the same grid order). Returns {(rid, q): error}."""
    import torch
    out = {}
    for r in res:
        if r.fields is None:
            continue
        for q in range(r.nrhs):
            y = ref.get((r.rid, q))
            if y is None:
                continue
            x = r.fields[q].reshape(-1)
            y = torch.as_tensor(y, device=x.device).reshape(-1)
            out[(r.rid, q)] = float(torch.linalg.norm(x - y) / torch.linalg.norm(y))
    return out


def _error_summary(errs, reference):
    import numpy as np
    v = np.array(list(errs.values()))
    if not len(v):
        return None
    return {"results": int(len(v)), "max": float(v.max()), "p99": float(np.percentile(v, 99)),
            "p50": float(np.percentile(v, 50)), "reference": reference}


def _sync_time(fn):
    import torch
    torch.cuda.synchronize()
    t = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t


def config_timings():
    """Wall times of the other workloads on one GPU (second of two runs each)."""
    import numpy as np
    import torch
    from fdfdpy_amd import configs as C
    from fdfdpy_amd.adjoint import FieldSolver
    from fdfdpy_amd.device import DeviceOperator
    from fdfdpy_amd.simulation import Simulation
    out = {}
    # 1: Ez 120x120, one point source, FP64 factor and solve through Simulation
    p = C.c1_point()
    for _ in range(2):
        sim = Simulation(p.omega, p.eps, p.dl, p.npml, 'Ez')
        sim.src = p.srcs[0]
        (_, _, ez), t = _sync_time(lambda: sim.solve_fields())
    out["c1_point_Ez_120"] = {"solve_fields_s": round(t, 4), "note": "assembly done at construction; "
                              "factor + refinement + secondary fields"}
    # 2: Hz 512x512, 8 point sources: c64 factor + FP64 refinement against the FP64 factor
    p = C.c2_design()
    b = torch.as_tensor(p.rhs(), device="cuda:0")
    res = {}
    for prec in ("c64", "fp64"):
        for _ in range(2):
            op = DeviceOperator(p.shape, p.omega, p.dl, p.npml, p.pol, p.L0, "cuda:0")
            op.set_eps(p.eps)
            _, tf = _sync_time(lambda: op.factor(prec))
            (x, rep_), ts = _sync_time(lambda: op.solve(b, precision=prec, tol=1e-10))
        res[prec] = (x, rep_, tf, ts)
    xm, xf = res["c64"][0], res["fp64"][0]
    err = [float(torch.linalg.vector_norm(xm[q] - xf[q]) / torch.linalg.vector_norm(xf[q])) for q in range(8)]
    out["c2_design_Hz_512_8rhs"] = {
        prec: {"factor_s": round(r[2], 4), "solve_8rhs_s": round(r[3], 4), "steps": r[1].iterations, "sweeps": r[1].sweeps,
               "max_rel_residual": r[1].max_rel_residual} for prec, r in res.items()}
    out["c2_design_Hz_512_8rhs"]["mixed_vs_fp64_max_rel_field_diff"] = max(err)
    # 4: Ez 300x300 adjoint gradient over the whole grid (forward + adjoint solve, resident factor)
    p = C.c4_adjoint()
    probe = torch.tensor(p.extra["probe"], device="cuda:0")
    src = torch.tensor(p.srcs[0], device="cuda:0")
    for _ in range(2):
        solver = FieldSolver(p.shape, p.omega, p.dl, p.npml, device="cuda:0")
        eps = torch.tensor(p.eps, device="cuda:0", requires_grad=True)

        def grad():
            Ez = solver(eps, src)
            (Ez.abs() ** 2 * probe).sum().backward()
            return eps.grad
        _, t = _sync_time(grad)
    out["c4_adjoint_Ez_300"] = {"objective_and_gradient_s": round(t, 4),
                                "note": "assembly, factor, forward solve, adjoint solve, gradient kernel"}
    # 5: Ez 400x400 Kerr, 8 amplitudes, to relative change 1e-10: the Born
    # iteration and Newton's method (the reference's default), each from the
    # linear field of a fresh simulation
    p = C.c5_kerr()
    per_amp = []
    for a in sorted(p.extra["amps"]):
        row = {"amp": round(float(a), 3)}
        fields = {}
        for method in ("born", "newton"):
            sim = Simulation(p.omega, p.eps, p.dl, p.npml, "Ez")
            sim.add_nl(p.extra["chi3"], p.extra["nl_region"])
            sim.src = p.srcs[0] * a
            (_, _, ez, conv), t = _sync_time(lambda: sim.solve_fields_nl(solver_nl=method, conv_threshold=1e-10,
                                                                           max_num_iter=50))
            fields[method] = ez
            row[method + "_steps"] = int(np.count_nonzero(conv))
            row[method + "_s"] = round(t, 4)
        row["newton_vs_born_max_rel_diff"] = float(np.linalg.norm(fields["newton"] - fields["born"]) /
                                                   np.linalg.norm(fields["born"]))
        per_amp.append(row)
    out["c5_kerr_Ez_400"] = per_amp
    return out


def run_stream(reqs, precision, slots, tol, engines=2, keep=False, fresh_cache=False):
    import torch
    from fdfdpy_amd.serving import ServingEngine, summarize
    # fresh_cache: the allocator's cached blocks go back to the device first
    # (outside the timed region). The latency legs take it: few requests are
    # in flight there, and blocks cut for the other precision's requests made
    # the allocator free and reallocate memory under load, which stalled the
    # FP64 leg for seconds. The throughput legs start from the warm cache: a
    # leg that starts empty allocates every request's buffers anew under full
    # load, which stalled about one run in three at its start.
    if fresh_cache:
        torch.cuda.empty_cache()
    eng = ServingEngine("cuda:0", precision=precision, slots=slots, tol=tol, engines=engines, keep_fields=keep)
    torch.cuda.synchronize()
    t0 = time.perf_counter()

    def progress(done, total):
        if done % 32 == 0 or done == total:
            log(f"  {precision}: {done}/{total} requests done, {time.perf_counter() - t0:.1f} s")
    res = eng.run(reqs, progress=progress)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    summ = summarize(res, wall)
    errs = sorted({r.error for r in res if r.error})
    if errs:
        log(f"  {precision}: {summ['failed']} failed requests; errors: {errs[:4]}")
    if eng.factor_engines:
        # batching of the factorizations: job-steps per shared step launch
        steps = sum(r.problem.shape[0] * (-(-r.problem.shape[1] // 32)) + 1 for r in reqs)
        launches = sum(e.launches for e in eng.factor_engines)
        summ["engine_launches"] = launches
        summ["jobs_per_launch"] = round(steps / launches, 3) if launches else None
    return summ, res


def _path_summary(summ):
    return {"solves_per_s": round(summ["solves_per_s"], 4),
            "requests_per_s": round(summ["requests_per_s"], 4),
            "p50_latency_ms": round(summ["p50_latency_s"] * 1e3, 3),
            "p99_latency_ms": round(summ["p99_latency_s"] * 1e3, 3),
            "p50_service_ms": round(summ["p50_service_s"] * 1e3, 3),
            "p99_service_ms": round(summ["p99_service_s"] * 1e3, 3),
            "max_rel_residual": summ["max_rel_residual"], "failed": summ["failed"],
            "fallbacks": summ["fallbacks"], "mean_steps": summ["mean_steps"], "mean_sweeps": summ["mean_sweeps"]}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is a design, not built)"}))
        sys.exit(2)
    if args.traffic_child:
        traffic_child()
        return
    import numpy as np
    from fdfdpy_amd import configs as C
    from fdfdpy_amd.serving import Request

    n_req = args.steps * args.batch
    problems = [C.c3_request(s) for s in range(n_req)]
    rng = np.random.default_rng(3)
    gaps = rng.exponential(1.0 / args.rate, size=n_req)
    arrivals = np.cumsum(gaps) - gaps[0]
    reqs = [Request(i, problems[i], float(arrivals[i])) for i in range(n_req)]

    # host-only legs first: their child processes start before this process
    # initializes the GPU
    traffic = None
    only = args.roofline_only or args.copy_calibration or args.configs or args.kernels_only or args.stream_only
    if not (args.no_traffic or only):
        log("traffic: two rocprofv3 --pmc passes over the roofline leg's launches")
        traffic = traffic_leg()
        log(f"traffic: {json.dumps(traffic)}")
    cpu, cpu_sols = None, {}
    if not (args.no_cpu_baseline or only):
        log("cpu: pool replay of a prefix of the stream")
        cpu, cpu_sols = cpu_baseline(reqs, args.cpu_seconds, args.cpu_workers)
        log(f"cpu: {json.dumps(cpu)}")

    import torch
    if args.roofline_only:
        print(json.dumps(roofline_leg()))
        return
    if args.copy_calibration:
        print(json.dumps(copy_calibration()))
        return
    if args.configs:
        print(json.dumps(config_timings()))
        return
    if args.kernels_only:
        print(json.dumps(kernels_only()))
        return

    log(f"stream: {n_req} requests, Poisson rate {args.rate}/s, {args.slots} slots")
    # warmup: separate requests of the same stream (8 per warmup step)
    for w in range(args.warmup):
        wreqs = [Request(i, C.c3_request(100000 + 8 * w + i), 0.0) for i in range(8)]
        for prec in (("c64",) if args.no_fp64 else ("c64", "fp64")):
            s, _ = run_stream(wreqs, prec, args.slots, args.tol, args.engines)
            log(f"warmup {w} {prec}: {s['solves']} solves in {s['wall_s']:.2f}s")

    # throughput leg: the whole stream, offered faster than it is served; the
    # results are kept (outside the timing) for the per-result errors
    mixed, res_m = run_stream(reqs, "c64", args.slots, args.tol, args.engines, keep=True)
    log(f"mixed: {json.dumps(mixed)}")
    fp64 = None
    errors = {}
    if cpu_sols:
        e = result_errors(res_m, cpu_sols)
        errors["vs_oracle"] = _error_summary(e, "the oracle's FP64 SuperLU solve of the same request and RHS "
                                                "(the CPU baseline's sample)")
        if errors["vs_oracle"]:
            errors["vs_oracle"]["per_result"] = {f"{rid}.{q}": float(f"{v:.3g}") for (rid, q), v in sorted(e.items())}
    if not args.no_fp64:
        fp64, res_f = run_stream(reqs, "fp64", args.slots, args.tol, args.engines, keep=True)
        log(f"fp64: {json.dumps(fp64)}")
        ref = {(r.rid, q): r.fields[q] for r in res_f if r.fields is not None for q in range(r.nrhs)}
        errors["vs_fp64_path"] = _error_summary(result_errors(res_m, ref),
                                                "the FP64-factor path's result of the same request and RHS")
        del res_f, ref
    del res_m
    if errors:
        log(f"errors: {json.dumps({k_: {a: b for a, b in v.items() if a != 'per_result'} for k_, v in errors.items() if v})}")
    if args.stream_only:
        print(json.dumps({"stream_c64": mixed, "stream_fp64": fp64}), flush=True)
        return
    # latency leg: a prefix of the stream at latency_load x the FP64 path's saturated request rate
    cap = (fp64 or mixed)["requests_per_s"]
    lat_rate = args.latency_load * cap
    k = min(args.latency_requests, n_req)
    lrng = np.random.default_rng(33)
    lgaps = lrng.exponential(1.0 / lat_rate, size=k)
    larr = np.cumsum(lgaps) - lgaps[0]
    lreqs = [Request(i, problems[i], float(larr[i])) for i in range(k)]
    log(f"latency leg: {k} requests at {lat_rate:.2f}/s")
    lat_m, _ = run_stream(lreqs, "c64", args.slots, args.tol, args.engines, fresh_cache=True)
    log(f"latency mixed: {json.dumps(lat_m)}")
    lat_f = None
    if not args.no_fp64:
        lat_f, _ = run_stream(lreqs, "fp64", args.slots, args.tol, args.engines, fresh_cache=True)
        log(f"latency fp64: {json.dumps(lat_f)}")
    roof = roofline_leg(traffic=traffic)
    log(f"roofline: {json.dumps(roof)}")

    ms_per_step = mixed["wall_s"] * 1e3 / args.steps
    out = {
        "metric": "solves/s + p50/p99 latency, FP32-LU+FP64-IR vs FP64 LU at residual 1e-10",
        "value": round(mixed["solves_per_s"], 4),
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 factor + f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"serving stream: {n_req} Ez requests, grids 128-768, Poisson arrivals "
                               f"{args.rate}/s, 1-4 RHS each, residual {args.tol:g}",
                   "requests": n_req, "batch": args.batch, "slots": args.slots, "engines": args.engines,
                   "latency_leg": f"first {k} requests, Poisson arrivals at {lat_rate:.2f}/s = "
                                  f"{args.latency_load} x the FP64 path's {cap:.2f} requests/s",
                   "parallelism": "single GPU"},
        "p50_latency_ms": round(lat_m["p50_latency_s"] * 1e3, 3),
        "p99_latency_ms": round(lat_m["p99_latency_s"] * 1e3, 3),
        "p50_service_ms": round(lat_m["p50_service_s"] * 1e3, 3),
        "p99_service_ms": round(lat_m["p99_service_s"] * 1e3, 3),
        "max_rel_residual": max(mixed["max_rel_residual"], lat_m["max_rel_residual"]),
        "fallbacks": mixed["fallbacks"], "failed": mixed["failed"] + lat_m["failed"],
        "mean_steps": mixed["mean_steps"], "mean_sweeps": mixed["mean_sweeps"],
        "error_per_result": errors or None,
        "saturated": _path_summary(mixed) | {"engine_launches": mixed.get("engine_launches"),
                                             "jobs_per_launch": mixed.get("jobs_per_launch")},
        "roofline": {k_: roof[k_] for k_ in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "roofline_detail": {k_: v for k_, v in roof.items()
                            if k_ not in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "cpu_baseline": cpu,
    }
    if fp64:
        out["fp64_path"] = {"saturated": _path_summary(fp64),
                            "latency_leg": _path_summary(lat_f) if lat_f else None}
        out["mixed_over_fp64"] = round(mixed["solves_per_s"] / fp64["solves_per_s"], 4)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
