"""Pins of the oracle's operator, solve and fields (linalg.py:41-149, simulation.py:112-178)."""

import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import (C_0, EPSILON_0, MU_0, OracleSimulation, construct_A, plane_wave_eigenvalue,
                    sfactor_vectors, solver_direct, stretch_symmetrizer)

OMEGA = 2 * np.pi * C_0 / 1.55e-6


def dense_stencil(pol, omega, eps, npml, dl, L0):
    """Dense A built cell by cell from the stretched 5-point stencil.

    Independent of scipy.sparse products: each row couples (i, j) to its
    periodic neighbours with the stretched second-difference weights.
    """
    Nx, Ny = eps.shape
    sxf, sxb, syf, syb = sfactor_vectors(omega, L0, Nx, Ny, npml, dl)
    e0, m0 = EPSILON_0 * L0, MU_0 * L0
    hx, hy = (Nx * dl) / Nx, (Ny * dl) / Ny
    A = np.zeros((Nx * Ny, Nx * Ny), complex)
    idx = lambda i, j: (i % Nx) * Ny + (j % Ny)
    for i in range(Nx):
        for j in range(Ny):
            p = idx(i, j)
            if pol == 'Ez':
                kx_p = 1 / (m0 * hx * hx * sxf[i] * sxb[(i + 1) % Nx])
                kx_m = 1 / (m0 * hx * hx * sxf[i] * sxb[i])
                ky_p = 1 / (m0 * hy * hy * syf[j] * syb[(j + 1) % Ny])
                ky_m = 1 / (m0 * hy * hy * syf[j] * syb[j])
                diag = omega ** 2 * e0 * eps[i, j]
            else:
                ex = lambda a, b: e0 * (eps[(a - 1) % Nx, b % Ny] + eps[a % Nx, b % Ny]) / 2
                ey = lambda a, b: e0 * (eps[a % Nx, (b - 1) % Ny] + eps[a % Nx, b % Ny]) / 2
                kx_p = 1 / (hx * hx * sxf[i] * ex(i + 1, j) * sxb[(i + 1) % Nx])
                kx_m = 1 / (hx * hx * sxf[i] * ex(i, j) * sxb[i])
                ky_p = 1 / (hy * hy * syf[j] * ey(i, j + 1) * syb[(j + 1) % Ny])
                ky_m = 1 / (hy * hy * syf[j] * ey(i, j) * syb[j])
                diag = omega ** 2 * m0
            A[p, idx(i + 1, j)] += kx_p
            A[p, idx(i - 1, j)] += kx_m
            A[p, idx(i, j + 1)] += ky_p
            A[p, idx(i, j - 1)] += ky_m
            A[p, p] += diag - kx_p - kx_m - ky_p - ky_m
    return A


def _A(pol, eps, npml, dl=0.02, L0=1e-6):
    N = eps.shape
    A, derivs = construct_A(OMEGA, [0, N[0] * dl], [0, N[1] * dl], eps, npml, pol, L0)
    return A, derivs


@pytest.mark.parametrize("pol", ["Ez", "Hz"])
def test_plane_wave_eigenvalues(pol):
    Nx, Ny, dl, epsv = 24, 18, 0.02, 2.25
    A, _ = _A(pol, np.full((Nx, Ny), epsv), [0, 0], dl)
    i, j = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing='ij')
    for a, b in [(0, 0), (1, 0), (3, 2), (-5, 7)]:
        kx, ky = 2 * np.pi * a / (Nx * dl), 2 * np.pi * b / (Ny * dl)
        v = np.exp(1j * (kx * i * dl + ky * j * dl)).reshape(-1)
        lam = plane_wave_eigenvalue(pol, OMEGA, 1e-6, dl, epsv, kx, ky)
        np.testing.assert_allclose(A @ v, lam * v, rtol=1e-10, atol=1e-10 * abs(lam))
        # a direct solve of a plane-wave RHS returns b / lambda exactly
        x = solver_direct(A, v)
        np.testing.assert_allclose(x, v / lam, rtol=1e-9)


@pytest.mark.parametrize("pol", ["Ez", "Hz"])
@pytest.mark.parametrize("npml", [[1, 1], [3, 2], [0, 2]])
def test_brute_force_assembly(pol, npml):
    rng = np.random.default_rng(11)
    eps = 1 + 11.25 * rng.random((9, 7))
    A, _ = _A(pol, eps, npml)
    D = dense_stencil(pol, OMEGA, eps, npml, 0.02, 1e-6)
    scale = np.abs(D).max()
    assert np.abs(A.toarray() - D).max() <= 1e-13 * scale


def test_tiny_dense_golden():
    g = np.load(os.path.join(GOLDEN, 'tiny_dense.npz'))
    for pol in ('Ez', 'Hz'):
        A, _ = construct_A(OMEGA, [0, 5 * 0.05], [0, 4 * 0.05], g['eps'], [1, 1], pol, 1e-6)
        np.testing.assert_allclose(A.toarray(), g[pol], rtol=0, atol=1e-14 * np.abs(g[pol]).max())


@pytest.mark.parametrize("pol", ["Ez", "Hz"])
def test_stretch_symmetrizer(pol):
    rng = np.random.default_rng(3)
    eps = 1 + 11.25 * rng.random((14, 11))
    npml = [4, 3]
    A, _ = _A(pol, eps, npml)
    d = stretch_symmetrizer(OMEGA, 1e-6, 14, 11, npml, 0.02)
    DA = (A.T.multiply(d).T).toarray()   # diag(d) @ A
    assert np.abs(DA - DA.T).max() <= 1e-13 * np.abs(DA).max()
    # a wrong stretch (the 'b' factors) breaks the symmetry
    _, sxb, _, syb = sfactor_vectors(OMEGA, 1e-6, 14, 11, npml, 0.02)
    wrong = (sxb[:, None] * syb[None, :]).reshape(-1)
    DW = (A.T.multiply(wrong).T).toarray()
    assert np.abs(DW - DW.T).max() > 1e-6 * np.abs(DW).max()


def test_hz_ez_uniform_relation():
    # uniform eps: (eps0' eps) A_Hz = mu0' A_Ez
    eps = np.full((10, 12), 3.0)
    Aez, _ = _A('Ez', eps, [2, 3])
    Ahz, _ = _A('Hz', eps, [2, 3])
    lhs = (EPSILON_0 * 1e-6 * 3.0) * Ahz.toarray()
    rhs = (MU_0 * 1e-6) * Aez.toarray()
    assert np.abs(lhs - rhs).max() <= 1e-13 * np.abs(rhs).max()


@pytest.mark.parametrize("pol", ["Ez", "Hz"])
def test_solve_residual_and_dense_agreement(pol):
    rng = np.random.default_rng(5)
    eps = 1 + 11.25 * (rng.random((16, 13)) < 0.5)
    A, _ = _A(pol, eps, [4, 4])
    b = rng.standard_normal(A.shape[0]) + 1j * rng.standard_normal(A.shape[0])
    x = solver_direct(A, b)
    assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b)
    xd = np.linalg.solve(A.toarray(), b)
    np.testing.assert_allclose(x, xd, rtol=1e-10, atol=1e-12 * np.abs(xd).max())


def test_zero_rhs_shortcut():
    A, _ = _A('Ez', np.ones((6, 6)), [1, 1])
    x = solver_direct(A, np.zeros(36))
    assert x.dtype == np.float64 and not x.any()
    with pytest.raises(ValueError):
        solver_direct(A, np.ones(36), solver='umfpack')


def test_fields_satisfy_ampere_law():
    # Ez: Dxf Hy - Dyf Hx - i omega eps0' eps Ez = src (from A Ez = i omega src)
    rng = np.random.default_rng(9)
    eps = 1 + 11.25 * (rng.random((30, 26)) < 0.5)
    sim = OracleSimulation(OMEGA, eps, 0.02, [6, 5], 'Ez')
    sim.src[15, 11] = 1
    sim.src[7, 20] = -0.5
    Hx, Hy, Ez = sim.solve_fields()
    D = sim.derivs
    lhs = D['Dxf'] @ Hy.reshape(-1) - D['Dyf'] @ Hx.reshape(-1) \
        - 1j * OMEGA * EPSILON_0 * 1e-6 * eps.reshape(-1) * Ez.reshape(-1)
    np.testing.assert_allclose(lhs, sim.src.reshape(-1), atol=1e-9)
    # Hz: the E fields give back -(Dxf Ey - Dyf Ex) - i omega mu0' Hz = src
    sim = OracleSimulation(OMEGA, eps, 0.02, [6, 5], 'Hz')
    sim.src[15, 11] = 1
    Ex, Ey, Hz = sim.solve_fields()
    D = sim.derivs
    curl = D['Dxf'] @ Ey.reshape(-1) - D['Dyf'] @ Ex.reshape(-1)
    lhs = -curl - 1j * OMEGA * MU_0 * 1e-6 * Hz.reshape(-1)
    np.testing.assert_allclose(lhs, sim.src.reshape(-1), atol=1e-9)


def test_c1_field_golden():
    g = np.load(os.path.join(GOLDEN, 'c1_field.npz'))
    eps = np.ones((120, 120))
    eps[45:75, 45:75] = 12.25
    sim = OracleSimulation(OMEGA, eps, 0.02, [15, 15], 'Ez')
    sim.src[60, 60] = 1
    _, _, Ez = sim.solve_fields()
    assert np.linalg.norm(Ez - g['Ez']) <= 1e-12 * np.linalg.norm(g['Ez'])


def test_notebook_modal_input_power():
    """notebooks/Examples.ipynb:105 holds W_in for the ridge-waveguide example
    (cell at :132-140); it exercises assembly, the direct solve, the H fields,
    the modal source and the flux probe end to end."""
    ref = json.load(open(os.path.join(GOLDEN, 'reference_values.json')))['notebooks/Examples.ipynb:105 W_in']
    omega = 2 * np.pi * 200e12
    eps = np.ones((600, 200))
    eps[:, 90:110] = 12.25
    sim = OracleSimulation(omega, eps, 0.01, [15, 15], 'Ez')
    sim.add_mode(3.5, 'x', [20, 100], 60, scale=10)
    sim.setup_modes()
    assert abs(sim.W_in / ref - 1) < 1e-9
