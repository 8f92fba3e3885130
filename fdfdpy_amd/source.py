"""Modal line sources (fdfdpy/source/mode.py:9-107).

A mode source injects the eigenmode of the 1D cross-section of the
permittivity along a line. The 1D eigenproblem is small (one line of cells)
and stays a host ARPACK shift-invert solve, as ``solver_eigs`` does
(linalg.py:110-119). The normalization run -- the full 2D solve of a straight
waveguide and the flux through the probe line -- goes through the device
solver (``Simulation.solve_fields``) like any other simulation.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl
import torch

from .constants import EPSILON_0, MU_0


def _host(a):
    return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def _periodic_difference(n, dl, kind):
    """1D periodic difference over n cells of size dl (derivatives.py:7-34 for
    one axis): forward 'f' is x[i+1] - x[i], backward 'b' is x[i] - x[i-1]."""
    i = np.arange(n)
    nb = (i + 1) % n if kind == 'f' else (i - 1) % n
    sgn = -1.0 if kind == 'f' else 1.0
    rows = np.concatenate([i, i])
    cols = np.concatenate([i, nb])
    vals = np.concatenate([np.full(n, sgn), np.full(n, -sgn)]) / dl
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def _line_slices(center, width, direction_normal):
    """Index ranges of the source line (mode.py:66-73)."""
    if direction_normal == "x":
        return [center[0], center[0] + 1], [int(center[1] - width / 2), int(center[1] + width / 2)]
    if direction_normal == "y":
        return [int(center[0] - width / 2), int(center[0] + width / 2)], [center[1], center[1] + 1]
    raise ValueError("The value of direction_normal is not x or y!")


class mode:
    """A modal source definition (mode.py:9-17); name and arguments as in
    the reference, so ``Simulation.add_mode`` / ``setup_modes`` read alike."""

    def __init__(self, neff, direction_normal, center, width, scale, order=1):
        self.neff = neff
        self.direction_normal = direction_normal
        self.center = center
        self.width = width
        self.order = order
        self.scale = scale

    def setup_src(self, simulation):
        """Input power from a straight-waveguide run, then the mode into
        ``simulation.src`` (mode.py:19-24)."""
        self.compute_normalization(simulation)
        self.insert_mode(simulation, simulation.src)

    def compute_normalization(self, simulation):
        """W_in and E2_in of the straight waveguide through the source line
        (mode.py:26-60). The waveguide keeps the cells of the source line
        whose eps exceeds 1, at the largest |eps| of the domain; the copy of
        the simulation keeps its source, and the mode is added to it."""
        eps_host = _host(simulation.eps_r)
        Nx, Ny = eps_host.shape
        core_eps = np.max(np.abs(eps_host))
        guide = np.ones((Nx, Ny))
        probe_at = list(self.center)
        if self.direction_normal == "x":
            # a guide along x: the source line's core cells, extended over every column
            guide[:, eps_host[self.center[0], :] > 1] = core_eps
            probe_at[0] = Nx - probe_at[0]
        elif self.direction_normal == "y":
            guide[eps_host[:, self.center[1]] > 1, :] = core_eps
            probe_at[1] = Ny - probe_at[1]
        else:
            raise ValueError("The value of direction_normal is not x or y!")

        straight = type(simulation)(simulation.omega, guide, simulation.dl, simulation.NPML, simulation.pol,
                                    simulation.L0, device=simulation.device)
        src = simulation.src
        straight.src = src.clone() if isinstance(src, torch.Tensor) else np.array(src, copy=True)
        self.insert_mode(straight, straight.src)
        straight.solve_fields()
        # the flux through the mirrored line at the far end of the guide
        simulation.W_in = straight.flux_probe(self.direction_normal, probe_at, self.width)
        # mode.py:59-60 reads 'Ez' whatever the polarization (a Hz mode has
        # no such field there either)
        e_line = _host(straight.fields['Ez'])
        simulation.E2_in = np.sum(np.square(np.abs(e_line)) * np.abs(_host(straight.src)))

    def insert_mode(self, simulation, destination):
        """Writes the scaled mode profile into ``destination`` on the source
        line (mode.py:62-107): the eigenvector of the 1D operator nearest to
        beta^2, beta = omega sqrt(mu0 eps0) neff, as |v| sign(Re v)."""
        eps0_ = EPSILON_0 * simulation.L0
        mu0_ = MU_0 * simulation.L0
        ix, iy = _line_slices(self.center, self.width, self.direction_normal)
        eps_line = _host(simulation.eps_r)[ix[0]:ix[1], iy[0]:iy[1]]
        N = eps_line.size
        Dxb = _periodic_difference(N, simulation.dl, 'b')
        Dxf = _periodic_difference(N, simulation.dl, 'f')
        # the permittivity on cell edges: average with the previous cell along
        # the first axis of the slice (linalg.py:16-22, as mode.py:82 calls it)
        eps_edge = (np.roll(eps_line, 1, axis=0) + eps_line) / 2
        T_eps = sp.diags(eps0_ * eps_line.reshape(-1), 0, shape=(N, N), format='csr')
        T_epsxinv = sp.diags(1 / (eps0_ * eps_edge.reshape(-1)), 0, shape=(N, N), format='csr')
        if simulation.pol == 'Ez':
            A = np.square(simulation.omega) * mu0_ * T_eps + Dxf @ Dxb
        elif simulation.pol == 'Hz':
            A = np.square(simulation.omega) * mu0_ * T_eps + T_eps @ Dxf @ T_epsxinv @ Dxb
        else:
            raise ValueError('Invalid polarization: {}'.format(str(simulation.pol)))
        beta = simulation.omega * np.sqrt(mu0_ * eps0_) * self.neff
        _, vecs = spl.eigs(A, k=self.order, sigma=np.square(beta), which='LM')
        v = vecs[:, self.order - 1] * self.scale
        profile = np.abs(v) * np.sign(np.real(v))
        shape = (1, -1) if self.direction_normal == 'x' else (-1, 1)
        profile = profile.reshape(shape)
        if isinstance(destination, torch.Tensor):
            destination[ix[0]:ix[1], iy[0]:iy[1]] = torch.as_tensor(profile, dtype=destination.dtype,
                                                                    device=destination.device)
        else:
            destination[ix[0]:ix[1], iy[0]:iy[1]] = profile
