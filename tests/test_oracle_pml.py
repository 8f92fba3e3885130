"""Pins of the oracle's PML and difference operators (pml.py:7-89, derivatives.py:7-34)."""

import os

import numpy as np
import pytest
from scipy.integrate import quad

from conftest import GOLDEN
from oracle import (C_0, EPSILON_0, ETA_0, MU_0, createDws, create_sfactor, sfactor_vectors,
                    sig_w, s_value)

OMEGA = 2 * np.pi * C_0 / 1.55e-6


def test_constants_match_si():
    # C_0 and ETA_0 derived from the two stored constants (constants.py:3-6)
    assert abs(C_0 - 299792458.0) / 299792458.0 < 1e-9
    assert abs(ETA_0 - 376.730313) < 1e-5
    assert EPSILON_0 == 8.85418782e-12 and MU_0 == 1.25663706e-6


def test_sigma_design_reflection():
    # sig_max is chosen so that 2 eta0 * integral_0^dw sigma = -lnR = 12
    for dw in (0.3, 0.4, 1.0):
        integral, _ = quad(lambda l: sig_w(l, dw), 0, dw)
        assert abs(2 * ETA_0 * integral - 12.0) < 1e-9
        # quartic grading
        assert abs(sig_w(dw / 2, dw) / sig_w(dw, dw) - 1 / 16) < 1e-14


def test_s_value_is_conductivity_stretch():
    l, dw = 0.2, 0.3
    s = s_value(l, dw, OMEGA, 1e-6)
    assert s.real == 1.0
    assert abs(-s.imag * OMEGA * EPSILON_0 * 1e-6 - sig_w(l, dw)) < 1e-12 * sig_w(l, dw)


@pytest.mark.parametrize("npml", [1, 15, 20, 25])
def test_sfactor_index_ranges(npml):
    Nw, hw = 120, 0.02
    f = create_sfactor([0, Nw * hw], 1e-6, 'f', OMEGA, Nw, npml)
    b = create_sfactor([0, Nw * hw], 1e-6, 'b', OMEGA, Nw, npml)
    nonunit_f = np.flatnonzero(f != 1)
    nonunit_b = np.flatnonzero(b != 1)
    # left layer holds Npml+1 cells; the right 'f' layer Npml-1 cells and the
    # right 'b' layer starts with an exact 1 (distance 0)
    left_f = nonunit_f[nonunit_f < Nw // 2]
    right_f = nonunit_f[nonunit_f >= Nw // 2]
    assert list(left_f) == list(range(0, npml + 1))
    assert list(right_f) == list(range(Nw - npml + 1, Nw))
    if npml >= 2:
        assert b[Nw - npml + 1] == 1
    assert list(nonunit_b[nonunit_b >= Nw // 2]) == list(range(Nw - npml + 2, Nw))
    # stretch grows toward the outer boundary
    assert np.all(np.diff(-f[:npml + 1].imag) < 0)
    if npml > 2:
        assert np.all(np.diff(-f[Nw - npml + 1:].imag) > 0)
    # distances on the half grid: ratios of (s - 1) follow (l1/l2)^4
    dw = npml * hw
    l_outer_b = (npml + 1) * hw          # 'b' outermost left cell lies beyond dw
    assert abs((b[0] - 1) / (s_value(dw, dw, OMEGA, 1e-6) - 1) - (l_outer_b / dw) ** 4) < 1e-12


def test_sfactor_no_pml_is_ones():
    assert np.all(create_sfactor([0, 2.0], 1e-6, 'f', OMEGA, 100, 0) == 1)


def test_sfactor_golden():
    g = np.load(os.path.join(GOLDEN, 'sfactors.npz'))
    for npml in (0, 1, 15, 20, 25):
        for kind in ('f', 'b'):
            v = create_sfactor([0, 120 * 0.02], 1e-6, kind, OMEGA, 120, npml)
            np.testing.assert_array_equal(v, g[f'n{npml}_{kind}'])


@pytest.mark.parametrize("N", [(7, 5), (12, 9)])
def test_difference_operators(N):
    d = [0.03, 0.05]
    for w in ('x', 'y'):
        Df = createDws(w, 'f', d, N).toarray()
        Db = createDws(w, 'b', d, N).toarray()
        np.testing.assert_allclose(Df, -Db.T, atol=0)
        np.testing.assert_allclose(Df.sum(axis=1), 0, atol=1e-12)
        np.testing.assert_allclose(Db.sum(axis=1), 0, atol=1e-12)
    # circulant spectrum along x: exp(i theta i) -> (e^{i theta} - 1)/dx
    Nx, Ny = N
    th = 2 * np.pi * 2 / Nx
    v = np.repeat(np.exp(1j * th * np.arange(Nx)), Ny)
    Df = createDws('x', 'f', d, N)
    np.testing.assert_allclose(Df @ v, (np.exp(1j * th) - 1) / d[0] * v, atol=1e-10)
    Db = createDws('x', 'b', d, N)
    np.testing.assert_allclose(Db @ v, (1 - np.exp(-1j * th)) / d[0] * v, atol=1e-10)


def test_sfactor_vectors_layout():
    sxf, sxb, syf, syb = sfactor_vectors(OMEGA, 1e-6, 40, 30, [5, 3], 0.02)
    assert sxf.shape == (40,) and syb.shape == (30,)
    assert np.all(sxf[6:36] == 1) and np.all(syf[4:27] == 1)
