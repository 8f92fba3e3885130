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

Also reported: the roofline of the dominant kernel (the Gauss-Jordan step
of the complex64 factor, timed with HIP events on its stream) and a CPU
baseline: the oracle's SuperLU restatement of the reference's 'scipy'
branch on a bounded sample of the same stream.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_MFMA_PEAK_TFLOPS = 157.3
FP64_MFMA_PEAK_TFLOPS = 78.6
# HBM bytes per k_gj_step<float, false> launch at n = 512 with symmetric steps
# (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 KiB, calibrated on a 256 MiB copy),
# from the rocprofv3 --pmc runs of `bench.py --roofline-only` summarized in
# profiles/k_gj_step_c64_512_roofline.md
GJ_STEP_C64_512_TRAFFIC_BYTES = 3.461e6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed batches of the stream")
    ap.add_argument("--warmup", type=int, default=1, help="untimed batches (of 8 requests each)")
    ap.add_argument("--batch", type=int, default=64, help="requests per step")
    ap.add_argument("--rate", type=float, default=200.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--slots", type=int, default=32, help="requests in flight")
    ap.add_argument("--engines", type=int, default=2, help="factorization engines (streams of step launches)")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="budget of the timed CPU sample (sizes beyond it are extrapolated)")
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


def roofline_leg(n=512, precision="c64"):
    """Factor one n x n operator with the stepwise kernel the request loop uses.

    The factorization is n*(n/32) back-to-back k_gj_step launches (+1 for
    the first pivot), timed with HIP events on their stream; the x-wrap
    correction that follows them in fdfd_factor is timed alone and taken out.
    The single persistent launch (k_gj_persist, the single-request path) is
    timed too.
    """
    from fdfdpy_amd import configs as C
    from fdfdpy_amd.device import DeviceOperator
    p = C.c3_request(7, n=n)
    op = DeviceOperator(p.shape, p.omega, p.dl, p.npml, p.pol, p.L0, "cuda:0")
    op.set_eps(p.eps)
    ms_all = _timed_factor(op, precision, stepwise=True)
    ms_wrap = _timed_wrap(op, precision)
    ms = ms_all - ms_wrap
    ms_persist = _timed_factor(op, precision, stepwise=False) - ms_wrap
    nt = (n + 31) // 32
    launches = n * nt + 1
    flops = 8.0 * n ** 4                       # 8 m n^3 real flops, m = n blocks
    achieved = flops / (ms * 1e-3) / 1e12
    peak = FP32_MFMA_PEAK_TFLOPS if precision == "c64" else FP64_MFMA_PEAK_TFLOPS
    tname = "float" if precision == "c64" else "double"
    return {
        "kernel": f"k_gj_step<{tname}, false>",
        "bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
        "frac": round(achieved / peak, 5),
        "traffic": GJ_STEP_C64_512_TRAFFIC_BYTES if (n, precision) == (512, "c64") else None,
        "workload": f"complex factor of a {n}x{n} Ez operator: {launches} launches, "
                    f"8*n^4 = {flops:.3e} flops per factor, 8*n^2*32 = {8.0 * n * n * 32:.3e} per launch",
        "avg_launch_us": round(ms * 1e3 / launches, 3), "factor_ms": round(ms, 3),
        "wrap_ms": round(ms_wrap, 3),
        "persistent_factor_ms": round(ms_persist, 3),
        "persistent_achieved_tflops": round(flops / (ms_persist * 1e-3) / 1e12, 3),
    }


def copy_calibration(nbytes=256 << 20, reps=3):
    """A plain device copy of known size (scales the PMC byte counters)."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"copy_bytes_each_way": nbytes, "copies": reps}


def cpu_baseline(reqs, seconds):
    """The oracle's restatement of the reference 'scipy' path (CSR assembly +
    SuperLU spsolve, refactored per right-hand side as linalg.py:122-149
    does), one host core, on a bounded sample of the stream: one solve of the
    first request of each grid size, smallest first, while the sample stays
    within `seconds`; sizes left out are extrapolated with the growth
    exponent of the two largest timed sizes. The stream's CPU rate is its
    solves divided by the sum of per-solve times over its requests."""
    import math
    from threadpoolctl import threadpool_limits
    from oracle import construct_A, solver_direct
    first = {}
    for r in reqs:
        first.setdefault(r.problem.shape[0], r.problem)
    sizes = sorted(first)
    per_solve, spent = {}, 0.0
    with threadpool_limits(limits=1):
        for n in sizes:
            if per_solve and spent >= seconds:
                break
            p = first[n]
            t = time.perf_counter()
            A, _ = construct_A(p.omega, [0, n * p.dl], [0, p.shape[1] * p.dl], p.eps, p.npml, p.pol, p.L0)
            solver_direct(A, p.srcs[0] * 1j * p.omega)
            per_solve[n] = time.perf_counter() - t
            spent += per_solve[n]
            log(f"  cpu: {n}x{n} solve {per_solve[n]:.2f} s")
    timed = sorted(per_solve)
    extrap = [n for n in sizes if n not in per_solve]
    if extrap:
        if len(timed) >= 2:
            a, b = timed[-2], timed[-1]
            expo = math.log(per_solve[b] / per_solve[a]) / math.log(b / a)
        else:
            expo = 3.0
        for n in extrap:
            per_solve[n] = per_solve[timed[-1]] * (n / timed[-1]) ** expo
    total = sum(r.problem.srcs.shape[0] * per_solve[r.problem.shape[0]] for r in reqs)
    nsolves = sum(r.problem.srcs.shape[0] for r in reqs)
    note = (f"; {extrap} extrapolated with exponent {expo:.2f} from {timed[-2:]}" if extrap else "")
    return {"value": round(nsolves / total, 5), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"one timed solve per grid size {timed} ({spent:.1f} s on 1 host core){note}; "
                      f"rate = {nsolves} stream solves / sum of per-solve times",
            "per_solve_s": {str(n): round(per_solve[n], 3) for n in sizes}}


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
        prec: {"factor_s": round(r[2], 4), "solve_8rhs_s": round(r[3], 4), "steps": r[1].iterations,
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


def run_stream(reqs, precision, slots, tol, engines=2):
    import torch
    from fdfdpy_amd.serving import ServingEngine, summarize
    eng = ServingEngine("cuda:0", precision=precision, slots=slots, tol=tol, engines=engines)
    torch.cuda.synchronize()
    t0 = time.perf_counter()

    def progress(done, total):
        if done % 32 == 0 or done == total:
            log(f"  {precision}: {done}/{total} requests done, {time.perf_counter() - t0:.1f} s")
    res = eng.run(reqs, progress=progress)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    summ = summarize(res, wall)
    if eng.factor_engines:
        # batching of the factorizations: job-steps per shared step launch
        steps = sum(r.problem.shape[0] * (-(-r.problem.shape[1] // 32)) + 1 for r in reqs)
        launches = sum(e.launches for e in eng.factor_engines)
        summ["engine_launches"] = launches
        summ["jobs_per_launch"] = round(steps / launches, 3) if launches else None
    return summ, res


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is a design, not built)"}))
        sys.exit(2)
    import numpy as np
    import torch
    from fdfdpy_amd import configs as C
    from fdfdpy_amd.serving import Request

    if args.roofline_only:
        print(json.dumps(roofline_leg()))
        return
    if args.copy_calibration:
        print(json.dumps(copy_calibration()))
        return
    if args.configs:
        print(json.dumps(config_timings()))
        return

    n_req = args.steps * args.batch
    log(f"stream: {n_req} requests, Poisson rate {args.rate}/s, {args.slots} slots")
    problems = [C.c3_request(s) for s in range(n_req)]
    rng = np.random.default_rng(3)
    gaps = rng.exponential(1.0 / args.rate, size=n_req)
    arrivals = np.cumsum(gaps) - gaps[0]
    reqs = [Request(i, problems[i], float(arrivals[i])) for i in range(n_req)]

    # warmup: separate requests of the same stream (8 per warmup step)
    for w in range(args.warmup):
        wreqs = [Request(i, C.c3_request(100000 + 8 * w + i), 0.0) for i in range(8)]
        for prec in (("c64",) if args.no_fp64 else ("c64", "fp64")):
            s, _ = run_stream(wreqs, prec, args.slots, args.tol, args.engines)
            log(f"warmup {w} {prec}: {s['solves']} solves in {s['wall_s']:.2f}s")

    mixed, res_m = run_stream(reqs, "c64", args.slots, args.tol, args.engines)
    log(f"mixed: {json.dumps(mixed)}")
    fp64 = None
    if not args.no_fp64:
        fp64, res_f = run_stream(reqs, "fp64", args.slots, args.tol, args.engines)
        log(f"fp64: {json.dumps(fp64)}")
    roof = roofline_leg()
    log(f"roofline: {json.dumps(roof)}")
    cpu = None if args.no_cpu_baseline else cpu_baseline(reqs, args.cpu_seconds)
    if cpu:
        log(f"cpu: {json.dumps(cpu)}")

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
                   "parallelism": "single GPU"},
        "p50_latency_ms": round(mixed["p50_latency_s"] * 1e3, 3),
        "p99_latency_ms": round(mixed["p99_latency_s"] * 1e3, 3),
        "p50_service_ms": round(mixed["p50_service_s"] * 1e3, 3),
        "p99_service_ms": round(mixed["p99_service_s"] * 1e3, 3),
        "max_rel_residual": mixed["max_rel_residual"], "fallbacks": mixed["fallbacks"],
        "failed": mixed["failed"], "mean_sweeps": mixed["mean_sweeps"],
        "roofline": {k: roof[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "roofline_detail": {k: roof[k] for k in ("kernel", "workload", "avg_launch_us", "factor_ms",
                                                 "persistent_factor_ms", "persistent_achieved_tflops")},
        "cpu_baseline": cpu,
    }
    if fp64:
        out["fp64_path"] = {
            "solves_per_s": round(fp64["solves_per_s"], 4),
            "p50_latency_ms": round(fp64["p50_latency_s"] * 1e3, 3),
            "p99_latency_ms": round(fp64["p99_latency_s"] * 1e3, 3),
            "p50_service_ms": round(fp64["p50_service_s"] * 1e3, 3),
            "p99_service_ms": round(fp64["p99_service_s"] * 1e3, 3),
            "max_rel_residual": fp64["max_rel_residual"], "failed": fp64["failed"],
            "mean_sweeps": fp64["mean_sweeps"]}
        out["mixed_over_fp64"] = round(mixed["solves_per_s"] / fp64["solves_per_s"], 4)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
