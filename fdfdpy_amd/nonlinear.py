"""Kerr nonlinearity, the Born iteration and Newton's method on the device.

Mirrors fdfdpy/nonlinearity.py:6-37 (the Kerr model), Simulation.compute_nl
(simulation.py:57-68), born_solve (nonlinear_solvers.py:13-54) and
newton_solve (nonlinear_solvers.py:57-149). Each step rewrites the
operator's diagonal on the device (fdfd_kerr_update) and solves with it; the
resident factor of the linear operator serves as the refinement
preconditioner, and the operator is refactored only when that refinement
stops contracting.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from .constants import EPSILON_0


class Nonlinearity:
    """Kerr nonlinearity (nonlinearity.py:6-37).

    ``chi`` is the internal coefficient (Simulation.add_nl divides the SI
    chi3 by L0**2, simulation.py:72). eps_nl = 3 chi nl_region |e|^2, scaled
    by (eps_r - 1)/(eps_max - 1) when eps_scale is set.
    """

    def __init__(self, chi, nl_region, nl_type='kerr', eps_scale=False, eps_max=None):
        self.chi = chi
        self.nl_region = nl_region
        self.nl_type = nl_type
        self.eps_scale = eps_scale
        self.eps_max = eps_max
        if nl_type != 'kerr':
            raise AssertionError("Only 'kerr' type nonlinearity is supported")
        if eps_scale and eps_max is None:
            raise AssertionError("Must provide eps_max when eps_scale is True")

    def coefficient(self, eps_r):
        """Real array c with eps_nl = c |e|^2."""
        c = 3 * self.chi * np.asarray(self.nl_region, dtype=np.float64)
        if self.eps_scale:
            c = c * ((np.real(np.asarray(eps_r)) - 1) / (self.eps_max - 1))
        return c

    def eps_nl(self, e, eps_r):
        return self.coefficient(eps_r) * np.square(np.abs(e))

    def dnl_de(self, e, eps_r):
        return self.coefficient(eps_r) * np.conj(e)

    def dnl_deps(self, e, eps_r):
        """d eps_nl / d eps_r: 3 chi nl_region |e|^2 / (eps_max - 1) when
        eps_scale is set, else 0 (nonlinearity.py:24, :31)."""
        if not self.eps_scale:
            return np.zeros(np.shape(self.nl_region))
        return 3 * self.chi * np.asarray(self.nl_region, dtype=np.float64) * np.square(np.abs(e)) / (self.eps_max - 1)


def kerr_coefficient(sim):
    """Sum of the Kerr coefficients of all nonlinearities of sim (device tensor)."""
    eps = sim.eps_r.detach().cpu().numpy() if isinstance(sim.eps_r, torch.Tensor) else sim.eps_r
    c = np.zeros((sim.Nx, sim.Ny))
    for nli in sim.nonlinearity:
        c = c + nli.coefficient(eps)
    return torch.as_tensor(c, dtype=torch.float64, device=sim.device)


def store_nl_state(sim, e, enl=None):
    """The host attributes compute_nl leaves on the simulation
    (simulation.py:57-68): eps_nl, dnl_de, dnl_deps (arrays of the grid's
    shape, summed over the nonlinearities) and Anl, the diagonal CSR matrix
    omega^2 eps0 L0 eps_nl. enl, when given, is eps_nl as computed on the
    device for the same field."""
    eh = e.detach().cpu().numpy() if isinstance(e, torch.Tensor) else np.asarray(e)
    eps = sim.eps_r.detach().cpu().numpy() if isinstance(sim.eps_r, torch.Tensor) else np.asarray(sim.eps_r)
    shape = (sim.Nx, sim.Ny)
    eps_nl = np.zeros(shape)
    dnl_de = np.zeros(shape, dtype=np.complex128)
    dnl_deps = np.zeros(shape)
    for nli in sim.nonlinearity:
        eps_nl = eps_nl + nli.eps_nl(eh, eps)
        dnl_de = dnl_de + nli.dnl_de(eh, eps)
        dnl_deps = dnl_deps + nli.dnl_deps(eh, eps)
    if enl is not None:
        eps_nl = enl.detach().cpu().numpy() if isinstance(enl, torch.Tensor) else np.asarray(enl)
    M = sim.Nx * sim.Ny
    sim.eps_nl = eps_nl
    sim.dnl_de = dnl_de
    sim.dnl_deps = dnl_deps
    sim.Anl = sp.spdiags(sim.omega ** 2 * EPSILON_0 * sim.L0 * eps_nl.reshape((-1,)), 0, M, M, format='csr')


def compute_nl(sim, e):
    """eps_nl of field e and the nonlinear diagonal on the device (simulation.py:57-68)."""
    et = torch.as_tensor(e, device=sim.device).to(torch.complex128)
    enl = sim.operator.kerr_update(et, kerr_coefficient(sim))
    store_nl_state(sim, et, enl)
    return enl


BORN_INNER_TOL = 1e-8


def born_solve(sim, Estart=None, conv_threshold=1e-10, max_num_iter=50, averaging=True,
               precision='fp64', tol=1e-12):
    """Born fixed point E_{k+1} = (A + Anl(E_k))^-1 b (nonlinear_solvers.py:13-54).

    Each step is taken in correction form, E_{k+1} = E_k + (A + Anl(E_k))^-1 r
    with r = b - (A + Anl(E_k)) E_k, which is the same iterate. The
    correction is solved to BORN_INNER_TOL relative to ||r|| (or tol, if
    looser). Its error is that fraction of the step's change and reappears in
    the next step's change, which is smaller by the Born contraction factor
    (~0.02 for a weak Kerr term); 1e-8 keeps the per-step changes within
    ~1e-6 of exact solves while saving the refinement steps a from-zero solve
    to 1e-12 of ||b|| spends.
    """
    conv_array = np.zeros((max_num_iter, 1))
    op = sim.operator
    if Estart is None:
        if sim.fields['Ez'] is None:
            (_, _, Ez) = sim.solve_fields(precision='fp64' if precision == 'fp64' else 'mixed')
        else:
            Ez = sim.fields['Ez']
    else:
        Ez = Estart
    E = torch.as_tensor(Ez, device=sim.device).to(torch.complex128).clone()
    coeff = kerr_coefficient(sim)
    b = sim._src_tensor() * (1j * sim.omega)
    convergence = np.inf
    enl = None
    try:
        for istep in range(max_num_iter):
            Eprev = E
            enl = op.kerr_update(Eprev, coeff)
            if bool(torch.any(b != 0)):
                r = b - op.apply(Eprev)
                if bool(torch.any(r != 0)):
                    d, rep = op.solve(r, precision=precision, tol=max(tol, BORN_INNER_TOL))
                    sim.last_report = rep
                    E = Eprev + d
                else:
                    E = Eprev.clone()
            else:
                E = torch.zeros_like(b)
            nE = torch.linalg.vector_norm(E).item()
            convergence = torch.linalg.vector_norm(E - Eprev).item() / nE if nE > 0 else 0.0
            conv_array[istep] = convergence
            if convergence < conv_threshold:
                break
    finally:
        if enl is not None:
            # the state of the reference's last compute_nl(Eprev)
            store_nl_state(sim, Eprev, enl)
        op.restore_linear()
    if convergence > conv_threshold:
        print("the simulation did not converge, reached {}".format(convergence))
    Hx, Hy = op.fields(E)
    return sim._out(Hx), sim._out(Hy), sim._out(E), conv_array


NEWTON_INNER_TOL = 1e-11
NEWTON_MAX_INNER = 50


def _conj_coupled_solve(op, j12, f, precision, tol):
    """d with J11 d + diag(j12) conj(d) = f, J11 the operator as set: J11^-1
    on the device (resident factor + refinement) inside a refinement of the
    conjugate-coupled system to tol relative to ||f||. This is the system
    solver_complex2real (linalg.py:152-186) writes in real form; the outer
    steps contract by ||J11^-1 diag(j12)||, the Kerr term's weight."""
    nf = torch.linalg.vector_norm(f).item()
    d = torch.zeros_like(f)
    if nf == 0:
        return d
    r = f
    for _ in range(NEWTON_MAX_INNER):
        dd, rep = op.solve(r, precision=precision, tol=max(tol, 1e-12))
        d = d + dd
        r = f - op.apply(d) - j12 * torch.conj(d)
        if torch.linalg.vector_norm(r).item() <= tol * nf:
            break
    return d


def newton_solve(sim, Estart=None, conv_threshold=1e-10, max_num_iter=50, averaging=True,
                 precision='fp64', tol=1e-12):
    """Newton's method for (A + Anl(E)) E = b (nonlinear_solvers.py:57-149).

    With eps_nl = c |E|^2 (c real), the Jacobian of f(E) = (A + Anl(E)) E - b
    (nl_eq_and_jac, nonlinear_solvers.py:118-149) is J11 d + J12 conj(d) with
    J11 = A + diag(omega^2 eps0' 2 c |E|^2) -- the operator with the Kerr
    term doubled, set on the device by fdfd_kerr_update -- and J12 =
    diag(omega^2 eps0' c E^2). Each step factors J11, solves J d = f and
    takes E - d; the run ends with a solve of (A + Anl(E)) X = b for the
    fields, as the reference does.
    """
    conv_array = np.zeros((max_num_iter, 1))
    op = sim.operator
    if Estart is None:
        if sim.fields['Ez'] is None:
            (_, _, Ez) = sim.solve_fields(precision='fp64' if precision == 'fp64' else 'mixed')
        else:
            Ez = sim.fields['Ez']
    else:
        Ez = Estart
    E = torch.as_tensor(Ez, device=sim.device).to(torch.complex128).clone()
    coeff = kerr_coefficient(sim)
    w = sim.omega ** 2 * EPSILON_0 * sim.L0
    b = sim._src_tensor() * (1j * sim.omega)
    convergence = np.inf
    enl = None
    try:
        for istep in range(max_num_iter):
            Eprev = E
            op.kerr_update(Eprev, coeff)
            f = op.apply(Eprev) - b
            j12 = (w * coeff).to(torch.complex128) * Eprev * Eprev
            op.kerr_update(Eprev, 2 * coeff)
            # one factorization of J11 per step (the reference refactors its
            # Jacobian every step too); the inner solves then take a few refinement steps
            op.factor(precision, refresh=True)
            E = Eprev - _conj_coupled_solve(op, j12, f, precision, NEWTON_INNER_TOL)
            nE = torch.linalg.vector_norm(E).item()
            convergence = torch.linalg.vector_norm(E - Eprev).item() / nE if nE > 0 else 0.0
            conv_array[istep] = convergence
            if convergence < conv_threshold:
                break
        enl = op.kerr_update(E, coeff)
        if bool(torch.any(b != 0)):
            X, rep = op.solve(b, precision=precision, tol=tol)
            sim.last_report = rep
        else:
            X = torch.zeros_like(b)
    finally:
        if enl is not None:
            # the reference's final compute_nl(Ez) before its last solve
            store_nl_state(sim, E, enl)
        op.restore_linear()
        # the resident factor is a Jacobian's: the next linear solve factors A afresh
        op.drop_factor(precision)
    if convergence > conv_threshold:
        print("the simulation did not converge, reached {}".format(convergence))
    Hx, Hy = op.fields(X)
    return sim._out(Hx), sim._out(Hy), sim._out(X), conv_array
