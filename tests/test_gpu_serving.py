"""The serving loop on the GPU: each result is the same alone and in the mix."""

import numpy as np
import pytest
import torch

from fdfdpy_amd import configs as C
from fdfdpy_amd.serving import Request, ServingEngine, summarize
from oracle import construct_A

pytestmark = pytest.mark.gpu


def _reqs():
    sizes = [128, 192, 128, 256, 192, 128]
    return [Request(i, C.c3_request(40 + i, n=sizes[i]), 0.0) for i in range(len(sizes))]


@pytest.mark.parametrize("precision", ["c64", "fp64"])
def test_alone_equals_in_mix(precision):
    reqs = _reqs()
    mix = ServingEngine("cuda:0", precision=precision, slots=4, keep_fields=True).run(reqs)
    assert all(not r.error for r in mix), [r.error for r in mix]
    for target in (0, 3):
        alone = ServingEngine("cuda:0", precision=precision, slots=1, keep_fields=True).run([reqs[target]])[0]
        a = alone.fields.cpu().numpy()
        m = mix[target].fields.cpu().numpy()
        # same operator, same kernels, same order of operations: identical bits
        assert np.array_equal(a, m), np.abs(a - m).max()
    for r, req in zip(mix, reqs):
        p = req.problem
        N = p.shape
        A, _ = construct_A(p.omega, [0, N[0] * p.dl], [0, N[1] * p.dl], p.eps, p.npml, p.pol, p.L0)
        x = r.fields.cpu().numpy()
        b = p.rhs()
        for q in range(b.shape[0]):
            res = np.linalg.norm(A @ x[q].reshape(-1) - b[q].reshape(-1)) / np.linalg.norm(b[q])
            assert res <= 1.5e-10


def test_memory_budget_and_factor_cache():
    reqs = _reqs()[:3]
    dup = Request(99, reqs[1].problem, 0.0)       # same operator as request 1
    eng = ServingEngine("cuda:0", precision="c64", slots=2, keep_fields=True, cache_bytes=1 << 32)
    res = eng.run(reqs)
    res2 = eng.run([dup])
    assert res2[0].factor_reused and not res[1].factor_reused
    assert np.array_equal(res2[0].fields.cpu().numpy(), res[1].fields.cpu().numpy())
    s = summarize(res, max(r.done for r in res))
    assert s["failed"] == 0 and s["max_rel_residual"] <= 1e-10


def test_engine_factors_match_direct_factors():
    """Factorizations batched in shared step launches equal the one-by-one
    launch-per-step ones."""
    import threading
    from fdfdpy_amd.device import DeviceOperator, FactorEngine
    eng = FactorEngine("cuda:0")
    probs = [C.c3_request(60 + i, n=n) for i, n in enumerate((128, 192, 256, 160))]
    precs = ["c64", "fp64", "c64", "c64"]
    ops_a, ops_b = [], []
    for p in probs:
        for ops, e in ((ops_a, None), (ops_b, eng)):
            op = DeviceOperator(p.shape, p.omega, p.dl, p.npml, p.pol, p.L0, "cuda:0")
            op.set_eps(p.eps)
            op.engine = e
            ops.append(op)
    for op, pr in zip(ops_a, precs):
        op.factor(pr, stepwise=True)
    threads = [threading.Thread(target=op.factor, args=(pr,)) for op, pr in zip(ops_b, precs)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    for a, b, pr in zip(ops_a, ops_b, precs):
        assert torch.equal(a.factors[pr][0], b.factors[pr][0])
    assert eng.launches < sum(p.shape[0] * ((p.shape[1] + 31) // 32) for p in probs)
    eng.close()


def test_engine_batches_of_more_than_eight_jobs():
    """12 factorizations of mixed sizes (odd and even tile counts) in shared
    step launches: two rows of critical workgroups and tile ranges that span
    jobs on every XCD; each equals its one-by-one launch-per-step factor."""
    import threading
    from fdfdpy_amd.device import DeviceOperator, FactorEngine
    eng = FactorEngine("cuda:0")
    sizes = (128, 160, 192, 96, 224, 128, 96, 256, 160, 96, 192, 128)
    probs = [C.c3_request(80 + i, n=n) for i, n in enumerate(sizes)]
    ops_a, ops_b = [], []
    for p in probs:
        for ops, e in ((ops_a, None), (ops_b, eng)):
            op = DeviceOperator(p.shape, p.omega, p.dl, p.npml, p.pol, p.L0, "cuda:0")
            op.set_eps(p.eps)
            op.engine = e
            ops.append(op)
    for op in ops_a:
        op.factor("c64", stepwise=True)
    threads = [threading.Thread(target=op.factor, args=("c64",)) for op in ops_b]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    for a, b in zip(ops_a, ops_b):
        assert torch.equal(a.factors["c64"][0], b.factors["c64"][0])
    eng.close()


def test_large_requests_alone_equal_in_mix():
    """The stream's 384 and 512 grids in a mix with smaller ones (two
    factorization engines, c64 factors): each field is bit-identical to the
    same request served alone, and its residual on the oracle's CSR operator
    is within the tolerance."""
    sizes = [384, 512, 128, 256, 384]
    reqs = [Request(i, C.c3_request(200 + i, n=sizes[i]), 0.0) for i in range(len(sizes))]
    mix = ServingEngine("cuda:0", precision="c64", slots=4, keep_fields=True).run(reqs)
    assert all(not r.error for r in mix), [r.error for r in mix]
    for target in (0, 1):
        alone = ServingEngine("cuda:0", precision="c64", slots=1, keep_fields=True).run([reqs[target]])[0]
        assert np.array_equal(alone.fields.cpu().numpy(), mix[target].fields.cpu().numpy())
    for r, req in zip(mix, reqs):
        p = req.problem
        A, _ = construct_A(p.omega, [0, p.shape[0] * p.dl], [0, p.shape[1] * p.dl], p.eps, p.npml, p.pol, p.L0)
        x = r.fields.cpu().numpy()
        b = p.rhs()
        worst = max(np.linalg.norm(A @ x[q].reshape(-1) - b[q].reshape(-1)) / np.linalg.norm(b[q])
                    for q in range(b.shape[0]))
        print(f"request {r.rid} ({p.shape[0]}^2, {b.shape[0]} RHS): {r.iterations} refinement steps, "
              f"oracle-CSR residual {worst:.2e}")
        assert worst <= 1.5e-10


def test_engine_many_waiters_on_tiny_jobs():
    """32 host threads submit tiny factorizations (short jobs, so waiters
    often arrive while their job's last step is being enqueued) and wait on
    them at once, repeatedly: every wait returns a complete factor, equal to
    the one-by-one launch-per-step factor. The completion state of a job is
    set under the engine's lock only after its event is recorded; a waiter
    that saw it earlier would destroy an unrecorded event and read a factor
    still being written."""
    import threading
    from fdfdpy_amd.device import DeviceOperator, FactorEngine
    eng = FactorEngine("cuda:0")
    probs = [C.c3_request(400 + i, n=n) for i, n in enumerate((96, 128, 96, 160))]
    ref = {}
    for i, p in enumerate(probs):
        op = DeviceOperator(p.shape, p.omega, p.dl, p.npml, p.pol, p.L0, "cuda:0")
        op.set_eps(p.eps)
        ref[i] = op.factor("c64", stepwise=True).clone()
    errors = []

    def worker(t):
        try:
            for rep in range(6):
                i = (t + rep) % len(probs)
                p = probs[i]
                op = DeviceOperator(p.shape, p.omega, p.dl, p.npml, p.pol, p.L0, "cuda:0")
                op.set_eps(p.eps)
                op.engine = eng
                f = op.factor("c64")
                torch.cuda.current_stream().synchronize()
                if not torch.equal(f, ref[i]):
                    errors.append((t, rep, i))
        except Exception as e:  # reported below
            errors.append(repr(e))
    threads = [threading.Thread(target=worker, args=(t,)) for t in range(32)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    eng.close()
    assert not errors, errors[:5]
