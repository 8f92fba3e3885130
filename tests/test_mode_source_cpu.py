"""Modal source profile (fdfdpy/source/mode.py:62-107) against the oracle.

The 1D eigensolve runs on the host in both; ARPACK's eigenvector carries an
arbitrary phase, which |v| sign(Re v) turns into an arbitrary global sign.
"""

from types import SimpleNamespace

import numpy as np
import pytest

from fdfdpy_amd.source import mode
from oracle import OracleMode, OracleSimulation

OMEGA = 2 * np.pi * 200e12


@pytest.mark.parametrize("pol,direction", [("Ez", "x"), ("Ez", "y"), ("Hz", "x"), ("Hz", "y")])
def test_mode_profile_matches_oracle(pol, direction):
    eps = np.ones((80, 70))
    if direction == "x":
        eps[:, 30:40] = 12.25
        center, width = [15, 35], 40
    else:
        eps[35:45, :] = 12.25
        center, width = [40, 12], 40
    sim = SimpleNamespace(eps_r=eps, dl=0.01, L0=1e-6, omega=OMEGA, pol=pol)
    got = np.zeros(eps.shape)
    mode(3.5, direction, center, width, 10).insert_mode(sim, got)
    ref_sim = OracleSimulation(OMEGA, eps, 0.01, [10, 10], pol)
    ref = np.zeros(eps.shape)
    OracleMode(3.5, direction, center, width, 10).insert_mode(ref_sim, ref)
    assert np.count_nonzero(ref) >= width - 2
    sign = np.sign(np.sum(got * ref))
    assert np.abs(sign * got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_bad_direction_rejected():
    sim = SimpleNamespace(eps_r=np.ones((10, 10)), dl=0.01, L0=1e-6, omega=OMEGA, pol="Ez")
    with pytest.raises(ValueError):
        mode(3.5, "z", [5, 5], 4, 1).insert_mode(sim, np.zeros((10, 10)))
