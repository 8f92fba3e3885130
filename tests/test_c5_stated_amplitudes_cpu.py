"""The Kerr workload at its stated amplitudes (line source 1e8 and 1e9) on
the oracle: the Born iteration of nonlinear_solvers.py:13-54 does not
converge there, in the reference or here.

At amplitude 1 the linear field peaks at |E| = 5.24 in the Kerr square, so
eps_nl = 3 chi3 / L0^2 |E|^2 reaches 2.3e12 at 1e8 and 2.3e14 at 1e9: the
first Born step suppresses the field in the square, the next restores it,
and the iteration settles into a 2-cycle with relative changes 0.642 and
0.599. The device run of the same case (tests/test_gpu_configs.py) is checked
against these values (tests/golden/c5_stated_amplitudes.json, written by
tests/golden/make_golden.py from this oracle).
"""

import json
import os

import numpy as np

from conftest import GOLDEN
from fdfdpy_amd import configs as C
from oracle import OracleSimulation


def test_c5_stated_amplitude_born_is_a_two_cycle():
    p = C.c5_kerr()
    region = p.extra['nl_region']
    gold = json.load(open(os.path.join(GOLDEN, 'c5_stated_amplitudes.json')))
    sim = OracleSimulation(p.omega, p.eps, p.dl, p.npml, 'Ez')
    sim.src = p.srcs[0] * 1e8
    sim.add_nl(p.extra['chi3'], region)
    _, _, E = sim.solve_fields()
    enl = 3 * p.extra['chi3'] / p.L0 ** 2 * region * np.abs(E) ** 2
    assert 2e12 < enl.max() < 3e12
    np.testing.assert_allclose(enl.max(), gold['1e+08']['eps_nl_max_linear_field'], rtol=1e-9)
    # eps_nl scales with amp^2: 1e9 gives 100 times more
    np.testing.assert_allclose(gold['1e+09']['eps_nl_max_linear_field'], 100 * enl.max(), rtol=1e-9)
    _, _, _, conv = sim.born_solve(conv_threshold=1e-10, max_num_iter=3)
    c = conv.ravel()
    assert np.all(c > 0.5)                      # never near the 1e-10 threshold
    assert abs(c[2] - c[0]) < 1e-6 * c[0]        # E_3 = E_1: a 2-cycle
    np.testing.assert_allclose(c, gold['1e+08']['born_conv'][:3], rtol=1e-6)
