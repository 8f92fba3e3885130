"""Writes the golden fixtures of this directory from the oracle.

The oracle itself is pinned by tests/test_oracle_*.py (closed forms,
brute-force assembly, the reference-held W_in of notebooks/Examples.ipynb).
Run from the repository root:  python tests/golden/make_golden.py
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import (C_0, OracleSimulation, construct_A, create_sfactor)  # noqa: E402


def main():
    omega = 2 * np.pi * C_0 / 1.55e-6
    # s-factor vectors (pml.py:20-39) on a 120-cell axis, dl = 0.02
    sf = {}
    for npml in (0, 1, 15, 20, 25):
        for kind in ('f', 'b'):
            sf[f'n{npml}_{kind}'] = create_sfactor([0, 120 * 0.02], 1e-6, kind, omega, 120, npml)
    np.savez(os.path.join(HERE, 'sfactors.npz'), **sf)

    # dense operators of a tiny grid (construct_A, linalg.py:41-107)
    rng = np.random.default_rng(7)
    N = (5, 4)
    eps = 1 + 11.25 * rng.random(N)
    dense = {'eps': eps}
    for pol in ('Ez', 'Hz'):
        A, _ = construct_A(omega, [0, N[0] * 0.05], [0, N[1] * 0.05], eps, [1, 1], pol, 1e-6)
        dense[pol] = A.toarray()
    np.savez(os.path.join(HERE, 'tiny_dense.npz'), **dense)

    # config 1 field: Ez 120x120, NPML 15, centred 30x30 square, src[60,60]
    eps = np.ones((120, 120))
    eps[45:75, 45:75] = 12.25
    sim = OracleSimulation(omega, eps, 0.02, [15, 15], 'Ez')
    sim.src[60, 60] = 1
    Hx, Hy, Ez = sim.solve_fields()
    np.savez(os.path.join(HERE, 'c1_field.npz'), Ez=Ez)

    # small Born trajectory (nonlinear_solvers.py:13-54)
    N = (64, 64)
    eps = np.ones(N)
    eps[24:40, 24:40] = 12.25
    region = np.zeros(N)
    region[24:40, 24:40] = 1
    sim = OracleSimulation(omega, eps, 0.02, [10, 10], 'Ez')
    sim.src[12, 24:40] = 40.0
    sim.add_nl(2.8e-18, region)
    _, _, Ez, conv = sim.born_solve(conv_threshold=1e-10, max_num_iter=50)
    np.savez(os.path.join(HERE, 'born_small.npz'), Ez=Ez, conv=conv.ravel(), eps_nl=sim.eps_nl)

    with open(os.path.join(HERE, 'reference_values.json'), 'w') as f:
        json.dump({'notebooks/Examples.ipynb:105 W_in': 0.0016689050319357455,
                   'notebooks/Examples_nonlinear.ipynb:143': 2.7e-06}, f, indent=1)
    c5_stated()


def c5_stated(n_steps=4):
    """The Kerr workload at its stated line-source amplitudes (1e8 and 1e9):
    the largest eps_nl of the linear field and the first Born steps'
    relative changes (nonlinear_solvers.py:13-54). Ez 400 x 400, NPML 20,
    eps 12.25 and chi3 = 2.8e-18 on [160:240]^2, src[50, 170:230] = amp."""
    N = 400
    omega = 2 * np.pi * C_0 / 1.55e-6
    eps = np.ones((N, N))
    eps[160:240, 160:240] = 12.25
    region = np.zeros((N, N))
    region[160:240, 160:240] = 1
    out = {}
    for amp in (1e8, 1e9):
        sim = OracleSimulation(omega, eps, 0.02, [20, 20], 'Ez')
        sim.src[50, 170:230] = amp
        sim.add_nl(2.8e-18, region)
        _, _, E = sim.solve_fields()
        eps_nl_max = float(np.max(3 * 2.8e-18 / 1e-12 * region * np.abs(E) ** 2))
        _, _, _, conv = sim.born_solve(conv_threshold=1e-10, max_num_iter=n_steps)
        out['%g' % amp] = {'eps_nl_max_linear_field': eps_nl_max, 'born_conv': [float(c) for c in conv.ravel()]}
    with open(os.path.join(HERE, 'c5_stated_amplitudes.json'), 'w') as f:
        json.dump(out, f, indent=1)


if __name__ == '__main__':
    if sys.argv[1:] == ['c5']:
        c5_stated()
    else:
        main()
