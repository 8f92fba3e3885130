"""Admission and scheduling of the serving loop, with the device work replaced (CPU)."""

import threading
import time

import numpy as np

from fdfdpy_amd import configs as C
from fdfdpy_amd.device import SolveReport
from fdfdpy_amd.serving import FactorCache, Request, ServingEngine, request_bytes, summarize


class FakeDevice:
    def __init__(self, seconds=0.02):
        self.seconds = seconds
        self.lock = threading.Lock()
        self.active = 0
        self.peak = 0
        self.order = []

    def __call__(self, req, stream):
        with self.lock:
            self.active += 1
            self.peak = max(self.peak, self.active)
            self.order.append(req.rid)
        time.sleep(self.seconds)
        with self.lock:
            self.active -= 1
        return None, SolveReport("c64", 3, True, 1e-11, [1e-11]), False


def stream(n, sizes=(128, 192, 256)):
    return [Request(i, C.c3_request(i, n=sizes[i % len(sizes)]), 0.002 * i) for i in range(n)]


def test_slots_bound_concurrency_and_all_complete():
    fake = FakeDevice()
    eng = ServingEngine("cpu", slots=3, memory_budget=1 << 40, serve_fn=fake)
    reqs = stream(12)
    res = eng.run(reqs)
    assert [r.rid for r in res] == list(range(12))
    assert all(not r.error for r in res)
    assert fake.peak <= 3 and eng.max_in_flight <= 3 and fake.peak >= 2
    assert all(r.done >= r.admitted >= r.arrival - 1e-3 for r in res)
    s = summarize(res, max(r.done for r in res))
    assert s["solves"] == sum(r.problem.srcs.shape[0] for r in reqs) and s["failed"] == 0


def test_memory_budget_limits_admission():
    fake = FakeDevice()
    reqs = stream(6, sizes=(256,))
    need = request_bytes(reqs[0].problem, "c64")
    eng = ServingEngine("cpu", slots=4, memory_budget=int(1.5 * need), serve_fn=fake)
    res = eng.run(reqs)
    assert all(not r.error for r in res)
    assert fake.peak == 1 and eng.max_bytes_in_flight <= 1.5 * need


def test_oversized_request_rejected_not_hung():
    fake = FakeDevice()
    reqs = stream(3, sizes=(512,))
    eng = ServingEngine("cpu", slots=2, memory_budget=1000, serve_fn=fake)
    res = eng.run(reqs)
    assert all("memory budget" in r.error for r in res)


def test_arrivals_respected():
    fake = FakeDevice(0.001)
    reqs = [Request(i, C.c3_request(i, n=128), 0.05 * i) for i in range(4)]
    eng = ServingEngine("cpu", slots=4, memory_budget=1 << 40, serve_fn=fake)
    res = eng.run(reqs)
    for r in res:
        assert r.admitted >= r.arrival - 1e-3
    assert fake.order == [0, 1, 2, 3]


def test_request_bytes_scale_with_precision():
    p = C.c3_request(0, n=512)
    assert request_bytes(p, "fp64") > 1.8 * request_bytes(p, "c64") * 0.9
    assert request_bytes(p, "c64") > 512 * 2 * 512 * 512 * 4


def test_factor_cache_lru():
    c = FactorCache(100)
    c.put("a", "A", 40)
    c.put("b", "B", 40)
    assert c.get("a") == "A"
    c.put("c", "C", 40)            # evicts the least recently used ('b')
    assert c.get("b") is None and c.get("a") == "A" and c.get("c") == "C"
    c.put("d", "D", 200)           # larger than the cache: not kept
    assert c.get("d") is None


def test_default_budget_counts_the_caching_allocators_unused_blocks(monkeypatch):
    # 10 GiB free on the device and 30 GiB held unused by torch's caching
    # allocator (reserved 50, allocated 20): the budget is 85% of 40 GiB, not
    # of the 10 GiB mem_get_info reports
    import torch
    gib = 1 << 30
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda device=None: (10 * gib, 288 * gib))
    monkeypatch.setattr(torch.cuda, "memory_reserved", lambda device=None: 50 * gib)
    monkeypatch.setattr(torch.cuda, "memory_allocated", lambda device=None: 20 * gib)
    eng = ServingEngine("cpu", slots=2, serve_fn=FakeDevice())
    assert eng.memory_budget == int(0.85 * 40 * gib)
