"""Pins of the oracle's Born iteration (nonlinear_solvers.py:13-54) and adjoint gradient."""

import os

import numpy as np
import scipy.sparse as sp

from conftest import GOLDEN
from oracle import (C_0, EPSILON_0, OracleSimulation, adjoint_gradient, solver_direct,
                    stretch_symmetrizer)

OMEGA = 2 * np.pi * C_0 / 1.55e-6


def _kerr_sim(chi, amp=40.0):
    N = (64, 64)
    eps = np.ones(N)
    eps[24:40, 24:40] = 12.25
    region = np.zeros(N)
    region[24:40, 24:40] = 1
    sim = OracleSimulation(OMEGA, eps, 0.02, [10, 10], 'Ez')
    sim.src[12, 24:40] = amp
    sim.add_nl(chi, region)
    return sim


def test_born_chi_zero_is_linear():
    sim = _kerr_sim(0.0)
    _, _, Elin = sim.solve_fields()
    _, _, E, conv = sim.born_solve()
    assert conv[0, 0] == 0.0 and conv[1, 0] == 0.0   # first step already stationary
    np.testing.assert_array_equal(E, Elin)


def test_born_fixed_point_and_golden():
    g = np.load(os.path.join(GOLDEN, 'born_small.npz'))
    sim = _kerr_sim(2.8e-18)
    _, _, E, conv = sim.born_solve(conv_threshold=1e-10)
    np.testing.assert_allclose(conv.ravel(), g['conv'], rtol=1e-6, atol=1e-14)
    assert np.linalg.norm(E - g['Ez']) <= 1e-12 * np.linalg.norm(g['Ez'])
    # E solves (A + Anl(E)) E = b up to the Born tolerance
    sim.compute_nl(E)
    b = (sim.src * 1j * OMEGA).reshape(-1)
    r = (sim.A + sim.Anl) @ E.reshape(-1) - b
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(b)
    # the nonlinearity is not negligible in this case
    assert 1e-3 < np.abs(sim.eps_nl).max() < 1


def test_adjoint_gradient_finite_differences():
    rng = np.random.default_rng(2)
    N = (40, 36)
    eps = np.ones(N)
    eps[14:26, 12:24] = rng.uniform(1, 12.25, size=(12, 12))
    probe = np.zeros(N)
    probe[32, 14:22] = 1

    def objective(e):
        s = OracleSimulation(OMEGA, e, 0.02, [8, 8], 'Ez')
        s.src[6, 12:24] = 1
        _, _, Ez = s.solve_fields()
        return float(np.sum(np.abs(Ez) ** 2 * probe))

    sim = OracleSimulation(OMEGA, eps, 0.02, [8, 8], 'Ez')
    sim.src[6, 12:24] = 1
    J, grad = adjoint_gradient(sim, probe)
    assert abs(J - objective(eps)) < 1e-14 * J
    h = 1e-4
    for (i, j) in [(15, 13), (20, 20), (25, 23), (14, 18)]:
        ep, em = eps.copy(), eps.copy()
        ep[i, j] += h
        em[i, j] -= h
        fd = (objective(ep) - objective(em)) / (2 * h)
        assert abs(fd - grad[i, j]) <= 1e-5 * abs(fd) + 1e-12 * J


def test_transpose_through_symmetrizer():
    # A^T x = b solved as D A^-1 D^-1 b (the device path) equals the CSR transpose solve
    rng = np.random.default_rng(4)
    eps = 1 + 11.25 * (rng.random((20, 18)) < 0.5)
    for pol in ('Ez', 'Hz'):
        sim = OracleSimulation(OMEGA, eps, 0.02, [5, 4], pol)
        d = stretch_symmetrizer(OMEGA, 1e-6, 20, 18, [5, 4], 0.02)
        b = rng.standard_normal(360) + 1j * rng.standard_normal(360)
        x1 = solver_direct(sim.A.T.tocsr(), b)
        x2 = d * solver_direct(sim.A, b / d)
        np.testing.assert_allclose(x2, x1, rtol=1e-9, atol=1e-12 * np.abs(x1).max())
        lam = rng.standard_normal(360)
        # bilinear identity (A^T lam)^T x = lam^T (A x)
        ref = np.dot(lam, sim.A @ x1)
        assert abs(np.dot(sim.A.T @ lam, x1) - ref) < 1e-9 * abs(ref)


def test_kerr_diagonal_is_omega2_eps0():
    sim = _kerr_sim(2.8e-18)
    e = np.full((64, 64), 2.0 + 1.0j)
    sim.compute_nl(e)
    enl = 3 * 2.8e-18 / 1e-12 * 5.0
    assert abs(sim.eps_nl[30, 30] - enl) < 1e-12 * enl and sim.eps_nl[0, 0] == 0
    d = sim.Anl.diagonal().reshape(64, 64)
    assert abs(d[30, 30] - OMEGA ** 2 * EPSILON_0 * 1e-6 * enl) < 1e-9 * abs(d[30, 30])
    assert isinstance(sim.Anl, sp.csr_matrix)


def test_complex2real_solves_the_conjugate_coupled_system():
    """solver_complex2real (linalg.py:152-186) returns x with
    A11 x + A12 conj(x) = b; dropping the conjugate coupling, or coupling
    x instead of conj(x), fails that equation."""
    from oracle import solver_complex2real
    rng = np.random.default_rng(4)
    n = 60
    A11 = sp.random(n, n, density=0.1, random_state=1) + 1j * sp.random(n, n, density=0.1, random_state=2) \
        + sp.eye(n) * (4 + 1j)
    A12 = sp.diags(rng.standard_normal(n) + 1j * rng.standard_normal(n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = solver_complex2real(A11, A12, b)
    assert np.linalg.norm(A11 @ x + A12 @ np.conj(x) - b) <= 1e-12 * np.linalg.norm(b)
    x_plain = solver_direct(A11 + A12, b)
    assert np.linalg.norm(A11 @ x_plain + A12 @ np.conj(x_plain) - b) > 1e-3 * np.linalg.norm(b)
    assert not np.any(solver_complex2real(A11, A12, np.zeros(n)))


def test_newton_reaches_the_born_fixed_point_faster():
    """newton_solve (nonlinear_solvers.py:57-149) converges to the same
    fixed point (A + Anl(E)) E = b as Born, in fewer steps, with the
    quadratic tail of an exact Jacobian."""
    born = _kerr_sim(2.8e-18)
    _, _, Eb, conv_b = born.born_solve(conv_threshold=1e-12)
    newton = _kerr_sim(2.8e-18)
    _, _, En, conv_n = newton.newton_solve(conv_threshold=1e-12)
    assert np.linalg.norm(En - Eb) <= 1e-10 * np.linalg.norm(Eb)
    steps_b, steps_n = np.count_nonzero(conv_b), np.count_nonzero(conv_n)
    assert steps_n < steps_b
    c = conv_n.ravel()[:steps_n]
    # quadratic: each change is at most ~ the square of the one before (relative)
    assert any(c[k + 1] < 10 * c[k] ** 2 for k in range(steps_n - 1) if c[k] < 1e-2)
    newton.compute_nl(En)
    b = (newton.src * 1j * OMEGA).reshape(-1)
    r = (newton.A + newton.Anl) @ En.reshape(-1) - b
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b)
