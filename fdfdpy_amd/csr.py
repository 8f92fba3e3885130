"""Host CSR copies of the system matrix and its derivative operators.

fdfdpy keeps ``Simulation.A`` and ``Simulation.derivs`` as scipy CSR matrices
(simulation.py:82-87, from construct_A at linalg.py:41-107). Here the operator
lives on the GPU; these functions build the same matrices on the host, on
demand, for user code that reads them. ``A`` is copied out of the device
operator's own 5-point bands; the derivative matrices are built from the
library's stretch factors (``fdfd_sfactors``), so both carry exactly the
values the device solver uses.
"""

from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib as L


def _periodic_difference(n, h, kind):
    """(1/h) times the periodic forward ('f': x[i+1] - x[i]) or backward
    ('b': x[i] - x[i-1]) difference on n cells (derivatives.py:7-34)."""
    i = np.arange(n)
    nb = (i + 1) % n if kind == 'f' else (i - 1) % n
    sgn = -1.0 if kind == 'f' else 1.0
    vals = np.concatenate([np.full(n, sgn), np.full(n, -sgn)]) / h
    return sp.csr_matrix((vals, (np.concatenate([i, i]), np.concatenate([i, nb]))), shape=(n, n))


def stretch_factors(omega, shape, NPML, dl, L0):
    """The reciprocal stretch factors (1/sxf, 1/sxb, 1/syf, 1/syb) of a grid,
    complex128 vectors of Nx, Nx, Ny, Ny entries (pml.py:42-89)."""
    Nx, Ny = int(shape[0]), int(shape[1])
    g = L.FdfdGrid(Nx, Ny, int(NPML[0]), int(NPML[1]), 0, 0, float(omega), float(dl), float(L0))
    out = [np.zeros(Nx, np.complex128), np.zeros(Nx, np.complex128),
           np.zeros(Ny, np.complex128), np.zeros(Ny, np.complex128)]
    L.check(L.lib().fdfd_sfactors(ctypes.byref(g), *[o.ctypes.data for o in out]))
    return tuple(out)


def stretched_derivatives(omega, shape, NPML, dl, L0):
    """The derivs dict of construct_A (linalg.py:55-58, 79-82, 95-100):
    Dxf, Dxb, Dyf, Dyb = diag(1/s) times the periodic differences, CSR,
    flattened C-order (x slow, y fast)."""
    Nx, Ny = int(shape[0]), int(shape[1])
    hx, hy = (Nx * dl) / Nx, (Ny * dl) / Ny   # dL() of linalg.py:25-33
    isxf, isxb, isyf, isyb = stretch_factors(omega, shape, NPML, dl, L0)
    ex, ey = sp.eye(Nx, format='csr'), sp.eye(Ny, format='csr')
    rows_x = lambda s: sp.diags(np.repeat(s, Ny), 0, format='csr')
    rows_y = lambda s: sp.diags(np.tile(s, Nx), 0, format='csr')
    return {
        'Dyb': (rows_y(isyb) @ sp.kron(ex, _periodic_difference(Ny, hy, 'b'))).tocsr(),
        'Dxb': (rows_x(isxb) @ sp.kron(_periodic_difference(Nx, hx, 'b'), ey)).tocsr(),
        'Dxf': (rows_x(isxf) @ sp.kron(_periodic_difference(Nx, hx, 'f'), ey)).tocsr(),
        'Dyf': (rows_y(isyf) @ sp.kron(ex, _periodic_difference(Ny, hy, 'f'))).tocsr(),
    }


def operator_csr(op):
    """A of a DeviceOperator as a host CSR matrix in the caller's orientation
    (the linear operator: Kerr diagonal updates are not included, as fdfdpy
    keeps them in Anl)."""
    Nx, Ny = op.shape
    M = Nx * Ny
    bands = op.op.view(-1, 2)[: 5 * M].cpu().numpy()
    bands = (bands[:, 0] + 1j * bands[:, 1]).reshape(5, M)
    base = op.base_c.view(-1, 2).cpu().numpy()
    bands[0] = base[:, 0] + 1j * base[:, 1]
    # solver grid: x is the block axis (Ny cells per block when transposed)
    nx, ny = (Ny, Nx) if op.transposed else (Nx, Ny)
    i, j = np.divmod(np.arange(M), ny)
    cols = np.stack([i * ny + j, ((i - 1) % nx) * ny + j, ((i + 1) % nx) * ny + j,
                     i * ny + (j - 1) % ny, i * ny + (j + 1) % ny])
    rows = np.broadcast_to(np.arange(M), cols.shape)
    if op.transposed:
        # solver cell (i, j) is caller cell (j, i), flattened as j * Ny_caller + i
        user = lambda p: (p % ny) * nx + p // ny
        rows, cols = user(rows), user(cols)
    A = sp.coo_matrix((bands.reshape(-1), (np.asarray(rows).reshape(-1), cols.reshape(-1))), shape=(M, M))
    return A.tocsr()
