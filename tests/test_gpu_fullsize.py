"""Full-size workloads checked through size-independent properties.

The residual of every returned field is recomputed on the CPU with the
oracle's CSR operator (an assembly independent of the device's), and the
mixed-precision fields are compared with the FP64 ones.
"""

import numpy as np
import pytest
import scipy.sparse.linalg as spl
import torch

from fdfdpy_amd import configs as C
from fdfdpy_amd.device import DeviceOperator
from oracle import construct_A

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def oracle_A(p):
    N = p.shape
    A, _ = construct_A(p.omega, [0, N[0] * p.dl], [0, N[1] * p.dl], p.eps, p.npml, p.pol, p.L0)
    return A


def residuals(A, x, b):
    return [np.linalg.norm(A @ x[q].reshape(-1) - b[q].reshape(-1)) / np.linalg.norm(b[q]) for q in range(len(b))]


def test_c2_hz_512_mixed_precision():
    p = C.c2_design()
    op = DeviceOperator(p.shape, p.omega, p.dl, p.npml, p.pol, p.L0, DEV)
    op.set_eps(p.eps)
    b = p.rhs()
    xm, rep = op.solve(torch.tensor(b, device=DEV), precision='c64', tol=1e-10)
    assert rep.converged and not rep.fell_back
    x64, rep64 = op.solve(torch.tensor(b, device=DEV), precision='fp64', tol=1e-12)
    A = oracle_A(p)
    xm_h, x64_h = xm.cpu().numpy(), x64.cpu().numpy()
    assert max(residuals(A, xm_h, b)) <= 1.5e-10
    assert max(residuals(A, x64_h, b)) <= 1.5e-12
    # the error of each result against the oracle's FP64 direct solve: one
    # SuperLU factor of the reference's CSR operator, 8 right-hand sides
    lu = spl.splu(A.tocsc())
    errs_m, errs_64 = [], []
    for q in range(b.shape[0]):
        xr = lu.solve(b[q].reshape(-1)).reshape(p.shape)
        errs_m.append(np.linalg.norm(xm_h[q] - xr) / np.linalg.norm(xr))
        errs_64.append(np.linalg.norm(x64_h[q] - xr) / np.linalg.norm(xr))
    print(f"c2 mixed: {rep.iterations} refinement steps, residual {rep.max_rel_residual:.2e}; error vs the "
          f"oracle's SuperLU solve per RHS: mixed {['%.2e' % e for e in errs_m]}, "
          f"FP64 {['%.2e' % e for e in errs_64]}")
    assert max(errs_m) < 1e-8 and max(errs_64) < 1e-9


def test_largest_stream_grid_768():
    p = C.c3_request(0, n=768)
    op = DeviceOperator(p.shape, p.omega, p.dl, p.npml, p.pol, p.L0, DEV)
    op.set_eps(p.eps)
    b = p.rhs()
    x, rep = op.solve(torch.tensor(b, device=DEV), precision='c64', tol=1e-10)
    assert rep.converged
    assert max(residuals(oracle_A(p), x.cpu().numpy(), b)) <= 1.5e-10


@pytest.mark.parametrize("precision", ["c64", "fp64"])
def test_block_beyond_persistent_sweep_limit(precision):
    """A 784 x 800 grid: blocks of 784 cells (padded 800), beyond the
    persistent sweep's register tiles, take the per-block sweep launches."""
    p = C.c3_request(7, n=768)
    eps = np.ones((800, 784))
    eps[200:600, 192:592] = C.design_eps(np.random.default_rng(11), 400)
    src = np.zeros((1, 800, 784))
    src[0, 100, 390] = 1
    prob = C.Problem('wide', 'Ez', p.omega, eps, [25, 25], src)
    op = DeviceOperator(prob.shape, prob.omega, prob.dl, prob.npml, prob.pol, prob.L0, DEV)
    assert not op.transposed and op.grid.ny == 784
    op.set_eps(prob.eps)
    b = prob.rhs()
    x, rep = op.solve(torch.tensor(b, device=DEV), precision=precision, tol=1e-10)
    print(f"784-cell blocks, {precision}: {rep.iterations} refinement steps, residual {rep.max_rel_residual:.2e}")
    assert rep.converged and not rep.fell_back
    assert max(residuals(oracle_A(prob), x.cpu().numpy(), b)) <= 1.5e-10
