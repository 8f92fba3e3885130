"""Full-size workloads checked through size-independent properties.

The residual of every returned field is recomputed on the CPU with the
oracle's CSR operator (an assembly independent of the device's), and the
mixed-precision fields are compared with the FP64 ones.
"""

import numpy as np
import pytest
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
    err = np.linalg.norm(xm_h - x64_h) / np.linalg.norm(x64_h)
    print(f"c2 mixed: {rep.iterations} sweeps, residual {rep.max_rel_residual:.2e}, error vs FP64 {err:.2e}")
    assert err < 1e-8


def test_largest_stream_grid_768():
    p = C.c3_request(0, n=768)
    op = DeviceOperator(p.shape, p.omega, p.dl, p.npml, p.pol, p.L0, DEV)
    op.set_eps(p.eps)
    b = p.rhs()
    x, rep = op.solve(torch.tensor(b, device=DEV), precision='c64', tol=1e-10)
    assert rep.converged
    assert max(residuals(oracle_A(p), x.cpu().numpy(), b)) <= 1.5e-10
