"""The solve's fallback order on the GPU: complex64 factor -> FP64 factor ->
FP64 factor with partial pivoting inside the 32x32 pivot tiles.

The reference's solver_direct (fdfdpy/linalg.py:122-149) factors with
Pardiso's own pivoting and always returns a solution; here the default
factors use diagonal pivot tiles, and a solve whose refinement stalls, or
whose factor meets a zero or non-finite pivot, moves to the next factor of
the order. Each case below forces one step of that order and checks the
final field against the oracle's SuperLU solve.
"""

import numpy as np
import pytest
import torch

from fdfdpy_amd import _lib as L
from fdfdpy_amd import configs as C
from fdfdpy_amd.device import DeviceOperator
from oracle import EPSILON_0, construct_A, solver_direct

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _problem(Nx=64, Ny=48, npml=(8, 8)):
    eps = np.ones((Nx, Ny), dtype=complex)
    eps[24:40, 16:32] = 12.25
    src = np.zeros((Nx, Ny))
    src[30, 20] = 1
    return C.omega_of(1.55e-6), eps, src, list(npml)


def _oracle_x(omega, eps, src, npml, dl=0.02):
    Nx, Ny = eps.shape
    A, _ = construct_A(omega, [0, Nx * dl], [0, Ny * dl], eps, npml, 'Ez', 1e-6)
    return A, solver_direct(A, src * 1j * omega).reshape(Nx, Ny)


def _device_solve(omega, eps, src, npml, **kw):
    op = DeviceOperator(eps.shape, omega, 0.02, npml, 'Ez', 1e-6, DEV)
    op.set_eps(eps)
    b = torch.tensor(src * 1j * omega, device=DEV)
    return op.solve(b, **kw)


def test_c64_stall_falls_back_to_fp64():
    """tol 1e-13 in 3 refinement steps: the complex64 factor reaches 1e-11 to
    1e-12 in 3 steps and needs a fourth, the FP64 factor gets to about 4e-15
    in 3 (lagged truncated wrap correction: the first step leaves the wrap
    coupling, the second corrects it)."""
    omega, eps, src, npml = _problem()
    x, rep = _device_solve(omega, eps, src, npml, precision='c64', tol=1e-13, max_iter=3)
    print(f"c64 -> {rep.precision}: {rep.iterations} steps, residual {rep.max_rel_residual:.2e}")
    assert rep.converged and rep.fell_back and rep.precision == 'fp64' and not rep.pivoted
    _, xr = _oracle_x(omega, eps, src, npml)
    xh = x.cpu().numpy()
    assert np.linalg.norm(xh - xr) / np.linalg.norm(xr) <= 1e-10


def test_singular_pivot_block_leaves_complex64_for_fp64():
    """eps[0, 0] is set so that the leading 2 x 2 block of the first pivot
    tile is singular to rounding while A is not. The complex64 factor cannot
    refine it; the FP64 factor, whose 2 x 2 pivot amplifies rounding by about
    1e13 (still below 1 / eps), refines to the tolerance."""
    omega, eps, src, npml = _problem()
    Nx, Ny = eps.shape
    A, _ = construct_A(omega, [0, Nx * 0.02], [0, Ny * 0.02], eps, npml, 'Ez', 1e-6)
    A = A.tocsr()
    w = omega ** 2 * EPSILON_0 * 1e-6
    eps[0, 0] = 1 + (A[0, 1] * A[1, 0] / A[1, 1] - A[0, 0]) / w
    A2, xr = _oracle_x(omega, eps, src, npml)
    A2 = A2.tocsr()
    det = A2[0, 0] * A2[1, 1] - A2[0, 1] * A2[1, 0]
    assert abs(det) < 1e-12 * abs(A2[0, 0] * A2[1, 1])     # the planted singular minor
    x, rep = _device_solve(omega, eps, src, npml, precision='c64', tol=1e-12)
    print(f"singular 2x2 pivot: solved by {rep.precision} pivoted={rep.pivoted} in {rep.iterations} steps; "
          f"factors with a bad pivot: {rep.failed_factors}")
    assert rep.converged and rep.fell_back and rep.precision == 'fp64'
    xh = x.cpu().numpy()
    assert np.linalg.norm(xh - xr) / np.linalg.norm(xr) <= 1e-10


def test_zero_pivot_falls_back_to_pivoted_factor(monkeypatch):
    """The last step of the order: the diagonal-pivot factors report a zero
    or non-finite pivot (FDFD_ERR_NUMERIC, forced here from the host for
    every unpivoted factor), and the FP64 factor with partial pivoting inside
    the tiles solves the system to the oracle's field."""
    omega, eps, src, npml = _problem()
    real_factor = DeviceOperator.factor

    def factor(self, precision="fp64", refresh=False, pivot=False, stepwise=False):
        resident_pivoted = precision in self.factors and self.factors[precision][2]
        if not pivot and not resident_pivoted:     # building an unpivoted factor
            self.factors.pop(precision, None)
            raise L.FdfdError(L.FDFD_ERR_NUMERIC, "zero or non-finite pivot in the block factorization")
        return real_factor(self, precision, refresh=refresh, pivot=pivot, stepwise=stepwise)
    monkeypatch.setattr(DeviceOperator, "factor", factor)
    x, rep = _device_solve(omega, eps, src, npml, precision='c64', tol=1e-12)
    print(f"forced zero pivots: solved by {rep.precision} pivoted={rep.pivoted}, failed {rep.failed_factors}")
    assert rep.converged and rep.fell_back and rep.pivoted and rep.precision == 'fp64'
    assert rep.failed_factors == [('c64', False), ('fp64', False)]
    _, xr = _oracle_x(omega, eps, src, npml)
    xh = x.cpu().numpy()
    assert np.linalg.norm(xh - xr) / np.linalg.norm(xr) <= 1e-10
    # every factor failing is an error of the solve, not a silent field
    monkeypatch.setattr(DeviceOperator, "factor", lambda self, *a, **k: (_ for _ in ()).throw(
        L.FdfdError(L.FDFD_ERR_NUMERIC, "zero pivot")))
    with pytest.raises(L.FdfdError):
        _device_solve(omega, eps, src, npml, precision='c64', tol=1e-10)


def test_no_factor_converges_raises():
    """Nothing in the order reaches tol 1e-15 in one step: NotConverged, as
    an error of the solve, not a silent field."""
    omega, eps, src, npml = _problem()
    with pytest.raises(L.NotConverged):
        _device_solve(omega, eps, src, npml, precision='c64', tol=1e-15, max_iter=1)
