"""Host CSR derivative operators and stretch factors (Simulation.derivs) against
the oracle's construct_A / S_create (linalg.py:55-58, pml.py:42-89)."""

import numpy as np
import pytest

from fdfdpy_amd.csr import stretch_factors, stretched_derivatives
from oracle import construct_A, sfactor_vectors

OMEGA = 1.2152e15


@pytest.mark.parametrize("shape,npml,pol", [((30, 20), [5, 4], "Ez"), ((25, 40), [0, 6], "Hz"),
                                            ((12, 9), [3, 0], "Ez"), ((600, 200), [15, 15], "Hz")])
def test_derivatives_match_oracle(shape, npml, pol):
    eps = 1 + 5 * np.random.default_rng(0).random(shape)
    _, ref = construct_A(OMEGA, [0, shape[0] * 0.02], [0, shape[1] * 0.02], eps, npml, pol, 1e-6)
    got = stretched_derivatives(OMEGA, shape, npml, 0.02, 1e-6)
    for k in ('Dxf', 'Dxb', 'Dyf', 'Dyb'):
        assert got[k].shape == ref[k].shape
        assert abs(got[k] - ref[k]).max() <= 1e-14 * abs(ref[k]).max(), k


def test_stretch_factors_match_oracle():
    got = stretch_factors(OMEGA, (40, 30), [7, 5], 0.02, 1e-6)
    ref = sfactor_vectors(OMEGA, 1e-6, 40, 30, [7, 5], 0.02)
    for g, r in zip(got, ref):
        assert np.abs(g - 1 / r).max() <= 1e-15 * np.abs(1 / r).max()
    # no PML: no stretching
    assert all(np.all(v == 1) for v in stretch_factors(OMEGA, (8, 8), [0, 0], 0.02, 1e-6))
