"""Continuous-batching request loop for FDFD solves on one GPU.

No reference counterpart: fdfdpy solves one problem per call and frees the
factor (linalg.py:135-138). Here requests of differing (Nx, Ny, omega, eps)
arrive over time and are admitted while a slot and device memory for their
operator, factor and workspaces are free; admitted requests run concurrently,
each on its own HIP stream, so their kernels overlap on the device. Every
right-hand side of a request reuses its resident factor, and factors of
identical operators are kept in an LRU cache across requests.
"""

from __future__ import annotations

import hashlib
import threading
import time
from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from .device import DeviceOperator, FactorEngine


@dataclass
class Request:
    rid: int
    problem: object            # configs.Problem
    arrival: float = 0.0       # seconds after the start of the run


@dataclass
class Result:
    rid: int
    shape: tuple
    nrhs: int
    fields: object             # (nrhs, Nx, Ny) complex128 on the device (or None)
    arrival: float
    admitted: float
    done: float
    precision: str
    iterations: int
    max_rel_residual: float
    fell_back: bool
    factor_reused: bool
    error: str = ""
    sweep_restarts: int = 0    # solves redone with per-block sweep launches
    sweeps: int = 0            # block sweeps T^-1 r of the solve

    @property
    def latency(self):
        return self.done - self.arrival

    @property
    def service(self):
        return self.done - self.admitted


def operator_key(p):
    h = hashlib.sha1(np.ascontiguousarray(p.eps).tobytes()).hexdigest()
    return (p.shape, p.pol, float(p.omega), float(p.dl), float(p.L0), tuple(p.npml), h)


def request_bytes(p, precision, nrhs=None):
    """Device bytes a request holds while it runs (operator, factor, workspaces, vectors)."""
    lib = L.lib()
    Nx, Ny = p.shape
    g = L.FdfdGrid(Nx, Ny, p.npml[0], p.npml[1], L.POL[p.pol], 0, p.omega, p.dl, p.L0)
    import ctypes
    gp = ctypes.byref(g)
    k = p.srcs.shape[0] if nrhs is None else nrhs
    pr = L.PREC[precision]
    return int(lib.fdfd_op_bytes(gp) * 1.2 + lib.fdfd_factor_bytes(gp, pr)
               + lib.fdfd_factor_workspace_bytes(gp, pr) + lib.fdfd_solve_workspace_bytes(gp, pr, k)
               + 16 * Nx * Ny * (3 * k + 2))


class FactorCache:
    """LRU of resident operators (with their factors) keyed by operator identity."""

    def __init__(self, capacity_bytes):
        self.capacity = capacity_bytes
        self.items = OrderedDict()     # key -> (DeviceOperator, bytes)
        self.bytes = 0
        self.lock = threading.Lock()

    def get(self, key):
        with self.lock:
            if key in self.items:
                self.items.move_to_end(key)
                return self.items[key][0]
        return None

    def put(self, key, op, nbytes):
        with self.lock:
            if nbytes > self.capacity:
                return
            self.items[key] = (op, nbytes)
            self.items.move_to_end(key)
            self.bytes += nbytes
            while self.bytes > self.capacity and self.items:
                _, (_, b) = self.items.popitem(last=False)
                self.bytes -= b


class ServingEngine:
    """Admits requests by device memory and runs them concurrently.

    precision: 'c64' (complex64 factor + FP64 refinement, FP64 fallback) or
    'fp64'. slots: requests in flight at once (one host thread and one HIP
    stream each). memory_budget: device bytes for in-flight requests
    (default: 85% of the memory free at construction, counting the
    blocks torch's caching allocator holds unused). cache_bytes: device
    bytes for factors kept after their request completes. batch_factors:
    the factorizations of in-flight requests advance together in shared
    Gauss-Jordan step launches (the native engine) instead of one launch
    sequence per request stream. engines: number of native engines, each with
    its own scheduler thread and stream; slot s uses engine s % engines, so
    the step launches of different engines overlap on the device.
    """

    def __init__(self, device="cuda:0", precision="c64", slots=4, memory_budget=None,
                 tol=1e-10, max_iter=80, cache_bytes=0, keep_fields=False, serve_fn=None,
                 batch_factors=True, engines=2):
        self.device = torch.device(device)
        self.precision = precision
        self.slots = slots
        if memory_budget is None:
            # free device memory plus the blocks torch's caching allocator
            # holds unused (earlier work on this device may have left most of
            # the memory there, where mem_get_info does not count it)
            free, _ = torch.cuda.mem_get_info(self.device)
            cached = torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
            memory_budget = int(0.85 * (free + max(cached, 0)))
        self.memory_budget = memory_budget
        self.tol = tol
        self.max_iter = max_iter
        self.cache = FactorCache(cache_bytes)
        self.keep_fields = keep_fields
        # serve_fn(request, stream) -> (fields, SolveReport, reused) replaces the
        # device work (scheduler tests run it on the CPU)
        self.serve_fn = serve_fn or self._serve
        self.streams = ([torch.cuda.Stream(device=self.device) for _ in range(slots)]
                        if serve_fn is None else [None] * slots)
        self.max_in_flight = 0
        self.max_bytes_in_flight = 0
        # factorizations of in-flight requests share step launches
        self.factor_engines = ([FactorEngine(self.device) for _ in range(max(1, engines))]
                               if (batch_factors and serve_fn is None) else [])
        self.factor_engine = self.factor_engines[0] if self.factor_engines else None

    def _engine_for(self, stream):
        if not self.factor_engines:
            return None
        slot = next(i for i, st in enumerate(self.streams) if st is stream)
        return self.factor_engines[slot % len(self.factor_engines)]

    # ------------------------------------------------------------ one request
    def _serve(self, req, stream):
        p = req.problem
        key = operator_key(p)
        reused = False
        with torch.cuda.stream(stream):
            op = self.cache.get(key) if self.cache.capacity else None
            if op is not None and self.precision in op.factors:
                reused = True
            else:
                op = DeviceOperator(p.shape, p.omega, p.dl, p.npml, p.pol, p.L0, self.device)
                op.engine = self._engine_for(stream)
                op.set_eps(p.eps)
            b = torch.as_tensor(p.rhs(), device=self.device)
            x, rep = op.solve(b, precision=self.precision, tol=self.tol, max_iter=self.max_iter)
            stream.synchronize()
        if self.cache.capacity and not reused:
            nbytes = op.op_bytes() + sum(f[0].numel() for f in op.factors.values())
            self.cache.put(key, op, nbytes)
        return x, rep, reused

    # ------------------------------------------------------------------ run
    def run(self, requests, realtime=True, progress=None):
        """Serve requests (sorted by arrival); returns Results in request order.

        progress(done, total), if given, is called after each completion.
        """
        reqs = sorted(requests, key=lambda r: r.arrival)
        results = {}
        cond = threading.Condition()
        state = {"used": 0, "free_slots": list(range(self.slots)), "next": 0}
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        def worker(req, slot, nbytes, admitted):
            err = ""
            x, rep, reused = None, None, False
            try:
                x, rep, reused = self.serve_fn(req, self.streams[slot])
            except Exception as e:  # reported per request, the loop continues
                err = f"{type(e).__name__}: {e}"
            done = now()
            res = Result(req.rid, req.problem.shape, req.problem.srcs.shape[0],
                         x if self.keep_fields else None, req.arrival, admitted, done,
                         rep.precision if rep else self.precision, rep.iterations if rep else 0,
                         rep.max_rel_residual if rep else float("nan"), rep.fell_back if rep else False,
                         reused, err, rep.restarts if rep else 0, rep.sweeps if rep else 0)
            with cond:
                results[req.rid] = res
                state["used"] -= nbytes
                state["free_slots"].append(slot)
                cond.notify_all()
                if progress is not None:
                    progress(len(results), len(reqs))

        threads = []
        with cond:
            while state["next"] < len(reqs):
                req = reqs[state["next"]]
                wait = req.arrival - now() if realtime else 0.0
                if wait > 0:
                    cond.wait(timeout=wait)
                    continue
                need = request_bytes(req.problem, self.precision)
                if need > self.memory_budget:
                    results[req.rid] = Result(req.rid, req.problem.shape, req.problem.srcs.shape[0], None,
                                              req.arrival, now(), now(), self.precision, 0, float("nan"),
                                              False, False, "request larger than the memory budget")
                    state["next"] += 1
                    continue
                if not state["free_slots"] or state["used"] + need > self.memory_budget:
                    cond.wait(timeout=0.05)
                    continue
                slot = state["free_slots"].pop(0)
                state["used"] += need
                state["next"] += 1
                self.max_in_flight = max(self.max_in_flight, self.slots - len(state["free_slots"]))
                self.max_bytes_in_flight = max(self.max_bytes_in_flight, state["used"])
                t = threading.Thread(target=worker, args=(req, slot, need, now()), daemon=True)
                threads.append(t)
                t.start()
        for t in threads:
            t.join()
        return [results[r.rid] for r in requests]


def summarize(results, wall):
    ok = [r for r in results if not r.error]
    lat = np.array([r.latency for r in ok]) if ok else np.array([np.nan])
    svc = np.array([r.service for r in ok]) if ok else np.array([np.nan])
    nsolves = sum(r.nrhs for r in ok)
    return {
        "requests": len(results), "failed": len(results) - len(ok), "solves": nsolves,
        "wall_s": wall, "solves_per_s": nsolves / wall if wall > 0 else float("nan"),
        "requests_per_s": len(ok) / wall if wall > 0 else float("nan"),
        "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
        "p50_service_s": float(np.percentile(svc, 50)), "p99_service_s": float(np.percentile(svc, 99)),
        "max_rel_residual": float(max((r.max_rel_residual for r in ok), default=float("nan"))),
        "fallbacks": sum(r.fell_back for r in ok),
        "mean_steps": float(np.mean([r.iterations for r in ok])) if ok else float("nan"),
        "mean_sweeps": float(np.mean([r.sweeps for r in ok])) if ok else float("nan"),
        "sweep_restarts": sum(r.sweep_restarts for r in ok),
    }
