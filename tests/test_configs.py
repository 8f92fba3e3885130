"""Seeded workload generators."""

import numpy as np

from fdfdpy_amd import configs as C


def test_c1():
    p = C.c1_point()
    assert p.shape == (120, 120) and p.pol == 'Ez' and p.npml == [15, 15]
    assert p.eps[45:75, 45:75].min() == 12.25 and p.eps.sum() == 120 * 120 + 900 * 11.25
    assert p.srcs.shape == (1, 120, 120) and p.srcs[0, 60, 60] == 1 and p.srcs.sum() == 1
    assert abs(p.omega - 1.2152e15) / 1.2152e15 < 1e-4


def test_c2_design_and_sources():
    p = C.c2_design()
    assert p.shape == (512, 512) and p.pol == 'Hz' and p.srcs.shape[0] == 8
    d = p.eps[128:384, 128:384]
    assert 1.0 <= d.min() and d.max() <= 12.25 and 0.3 < (d.mean() - 1) / 11.25 < 0.7
    assert np.all(p.eps[:128] == 1)
    for q in range(8):
        (i,), (j,) = np.nonzero(p.srcs[q].sum(axis=1))[0:1], np.nonzero(p.srcs[q].sum(axis=0))[0:1]
        assert not (128 <= i < 384 and 128 <= j < 384)
        assert 20 < i < 492 and 20 < j < 492
    np.testing.assert_array_equal(C.c2_design().eps, p.eps)   # seeded


def test_c3_stream_mix():
    reqs, arr = C.c3_stream(64, rate=10.0)
    sizes = {r.shape[0] for r in reqs}
    assert sizes <= set(C.STREAM_SIZES) and len(sizes) >= 4
    for r in reqs:
        n = r.shape[0]
        assert r.npml == [C.stream_npml(n)] * 2 and 15 <= r.npml[0] <= 25
        assert 1 <= r.srcs.shape[0] <= 4
        assert 1.3e-6 <= r.extra['lambda'] <= 1.7e-6
    assert np.all(np.diff(arr) >= 0) and arr[0] == 0
    assert C.c3_request(5).eps.tobytes() == C.c3_request(5).eps.tobytes()


def test_c4_c5():
    p = C.c4_adjoint()
    assert p.shape == (300, 300) and len(p.extra['fd_pixels']) == 32
    assert all(p.extra['design'][i, j] for i, j in p.extra['fd_pixels'])
    assert p.srcs[0, 40, 130:170].sum() == 40 and p.extra['probe'].sum() == 40
    q = C.c5_kerr()
    assert q.shape == (400, 400) and len(q.extra['amps']) == 8
    assert np.all((q.extra['amps'] >= 3) & (q.extra['amps'] <= 60))


def test_point_sources_refuse_impossible_layouts():
    import pytest
    rng = np.random.default_rng(0)
    with pytest.raises(ValueError):
        C.point_sources(rng, 64, 14, 16, 48, 1)     # every PML-free cell is in the design region
    assert len(C.point_sources(rng, 128, 15, 32, 96, 4)) == 4
