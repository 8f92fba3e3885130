"""CPU restatement of fdfdpy's 2D FDFD solve path -- the oracle.

TEST INFRASTRUCTURE. This module is the checker for the HIP implementation in
``fdfdpy_amd``; it is never part of the product path. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import it.

It restates, with numpy/scipy.sparse, the reference's own ``solver='scipy'``
branch: CSR assembly of the SC-PML 5-point operator and a SuperLU direct solve
that refactors on every call. Each function cites the reference lines it
follows (paths relative to the fdfdpy repository root).

Pinning (what checks each function against something other than itself; the
tests live in ``tests/test_oracle_*.py``):

* constants            -- C_0 and ETA_0 against the SI values they imply.
* sig_w / s_value      -- the design reflection: 2*ETA_0*integral(sigma) = -lnR,
                          and the (l/dw)**m grading.
* create_sfactor       -- index ranges fixed by special cells (interior s == 1,
                          first right 'b' cell == 1, Npml < 1 -> ones), plus
                          the grading law across cells.
* createDws            -- dxf == -dxb.T, zero row sums, circulant spectrum.
* construct_A (Ez, Hz) -- plane-wave eigenvalues of the periodic operator
                          (closed form), a dense assembly from explicit
                          per-cell stencil loops on tiny grids, and the
                          symmetry of diag(sxf*syf) @ A.
* solver_direct        -- residual and agreement with numpy.linalg.solve.
* solve_fields         -- H fields against the discrete Faraday law and the
                          reference-held modal input power of the ridge
                          waveguide example (W_in = 0.0016689050319357455).
* born_solve           -- chi = 0 reduces to the linear solve; the returned
                          field is a fixed point of (A + Anl(E)) E = b.
* adjoint_gradient     -- central finite differences of the objective.
* solver_complex2real  -- its solution satisfies A11 x + A12 conj(x) = b;
                          dropping the conjugate coupling fails that.
* newton_solve         -- reaches Born's fixed point (A + Anl(E)) E = b, in
                          fewer steps, with a quadratic tail.
"""

from __future__ import annotations

import copy
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

__all__ = [
    "EPSILON_0", "MU_0", "C_0", "ETA_0", "DEFAULT_LENGTH_SCALE",
    "sig_w", "s_value", "create_sfactor", "sfactor_vectors", "S_create",
    "createDws", "grid_average", "dL", "construct_A", "solver_direct",
    "stretch_symmetrizer", "solver_complex2real", "OracleSimulation", "OracleMode",
    "adjoint_gradient", "plane_wave_eigenvalue",
]

# constants.py:3-6 (values are part of parity) ------------------------------
EPSILON_0 = 8.85418782e-12
MU_0 = 1.25663706e-6
C_0 = np.sqrt(1.0 / EPSILON_0 / MU_0)
ETA_0 = np.sqrt(MU_0 / EPSILON_0)
DEFAULT_LENGTH_SCALE = 1e-6  # constants.py:10


# pml.py:7-17 ---------------------------------------------------------------
def sig_w(l, dw, m=4, lnR=-12):
    """Polynomial conductivity grading (pml.py:7-11)."""
    sig_max = -(m + 1) * lnR / (2 * ETA_0 * dw)
    return sig_max * (np.asarray(l, dtype=np.float64) / dw) ** m


def s_value(l, dw, omega, L0):
    """Complex stretch factor s = 1 - i sigma/(omega eps0 L0) (pml.py:14-17)."""
    return 1 - 1j * sig_w(l, dw) / (omega * EPSILON_0 * L0)


def create_sfactor(wrange, L0, s, omega, Nw, Nw_pml):
    """Stretch factors along one axis for the 'f' or 'b' grid (pml.py:20-39).

    The left layer covers cells [0, Nw_pml] and the right layer cells
    [Nw - Nw_pml + 1, Nw - 1]; the distances differ between 'f' and 'b'.
    The left test wins where the two ranges overlap (pml.py:29-38).
    """
    out = np.ones(Nw, dtype=np.complex128)
    if Nw_pml < 1:
        return out
    hw = np.diff(wrange)[0] / Nw
    dw = Nw_pml * hw
    i = np.arange(Nw)
    left = i <= Nw_pml
    right = (~left) & (i > Nw - Nw_pml)
    if s == 'f':
        l_left = hw * (Nw_pml - i + 0.5)
        l_right = hw * (i - (Nw - Nw_pml) - 0.5)
    elif s == 'b':
        l_left = hw * (Nw_pml - i + 1)
        l_right = hw * (i - (Nw - Nw_pml) - 1)
    else:
        return out
    out[left] = s_value(l_left[left], dw, omega, L0)
    out[right] = s_value(l_right[right], dw, omega, L0)
    return out


def sfactor_vectors(omega, L0, Nx, Ny, NPML, dl):
    """The four s-value vectors (sxf, sxb, syf, syb) of a grid (pml.py:68-71)."""
    xrange = [0, float(Nx * dl)]
    yrange = [0, float(Ny * dl)]
    return (create_sfactor(xrange, L0, 'f', omega, Nx, NPML[0]),
            create_sfactor(xrange, L0, 'b', omega, Nx, NPML[0]),
            create_sfactor(yrange, L0, 'f', omega, Ny, NPML[1]),
            create_sfactor(yrange, L0, 'b', omega, Ny, NPML[1]))


def S_create(omega, L0, N, Npml, xrange, yrange):
    """Diagonal 1/s matrices (Sxf, Sxb, Syf, Syb), C-order (pml.py:42-89)."""
    Nx, Ny = int(N[0]), int(N[1])
    sxf = create_sfactor(xrange, L0, 'f', omega, Nx, Npml[0])
    sxb = create_sfactor(xrange, L0, 'b', omega, Nx, Npml[0])
    syf = create_sfactor(yrange, L0, 'f', omega, Ny, Npml[1])
    syb = create_sfactor(yrange, L0, 'b', omega, Ny, Npml[1])
    # x varies along axis 0 (rows of the 2D array), y along axis 1
    diag = lambda v: sp.diags(v, 0, shape=(Nx * Ny, Nx * Ny), format='csr')
    return (diag(np.repeat(1 / sxf, Ny)), diag(np.repeat(1 / sxb, Ny)),
            diag(np.tile(1 / syf, Nx)), diag(np.tile(1 / syb, Nx)))


# derivatives.py:7-34 -------------------------------------------------------
def createDws(w, s, dL_, N):
    """Periodic forward/backward difference operators (derivatives.py:7-34).

    The 1D forward difference is (x[i+1] - x[i]) with x[Nw] == x[0], the
    backward difference (x[i] - x[i-1]) with x[-1] == x[Nw-1]. x is the slow
    (block) index of the C-order flattening, y the fast one.
    """
    Nx = int(N[0])
    Ny = int(N[1]) if len(N) > 1 else 1
    dx = dL_[0]
    dy = dL_[1] if len(N) > 1 else np.inf

    def d1(n, kind):
        if kind == 'f':
            return sp.diags([-1, 1, 1], [0, 1, -n + 1], shape=(n, n))
        return sp.diags([1, -1, -1], [0, -1, n - 1], shape=(n, n))

    if w == 'x':
        return (1 / dx) * sp.kron(d1(Nx, s), sp.eye(Ny), format='csr')
    return (1 / dy) * sp.kron(sp.eye(Nx), d1(Ny, s), format='csr')


# linalg.py:16-33 -----------------------------------------------------------
def grid_average(center_array, w):
    """Average with the previous cell along x or y, periodic (linalg.py:16-22)."""
    axis = {'x': 0, 'y': 1}[w]
    return (np.roll(center_array, 1, axis=axis) + center_array) / 2


def dL(N, xrange, yrange=None):
    """Cell sizes from the domain extent (linalg.py:25-33)."""
    if yrange is None:
        L = np.array([np.diff(xrange)[0]])
    else:
        L = np.array([np.diff(xrange)[0], np.diff(yrange)[0]])
    return L / N


# linalg.py:41-107 ----------------------------------------------------------
def construct_A(omega, xrange, yrange, eps_r, NPML, pol, L0, averaging=True):
    """System matrix and stretched derivatives (linalg.py:41-107).

    Ez: A = Dxf Dxb / mu0' + Dyf Dyb / mu0' + omega^2 eps0' diag(eps)
    Hz: A = Dxf Tx^-1 Dxb + Dyf Ty^-1 Dyb + omega^2 mu0' I, where the reference
        writes ``eye(M)`` with ``eye`` undefined (linalg.py:93, a NameError);
        this restatement uses the evident identity.
    """
    N = np.asarray(eps_r.shape)
    M = int(np.prod(N))
    eps0_ = EPSILON_0 * L0
    mu0_ = MU_0 * L0
    if pol not in ('Ez', 'Hz'):
        raise ValueError("something went wrong and pol is not one of Ez, Hz, "
                         "instead was given {}".format(pol))
    Sxf, Sxb, Syf, Syb = S_create(omega, L0, N, NPML, xrange, yrange)
    d = dL(N, xrange, yrange)
    Dyb = Syb @ createDws('y', 'b', d, N)
    Dxb = Sxb @ createDws('x', 'b', d, N)
    Dxf = Sxf @ createDws('x', 'f', d, N)
    Dyf = Syf @ createDws('y', 'f', d, N)
    if pol == 'Ez':
        Teps = sp.diags(eps0_ * eps_r.reshape(-1), 0, shape=(M, M), format='csr')
        A = (Dxf / mu0_) @ Dxb + (Dyf / mu0_) @ Dyb + omega ** 2 * Teps
    else:
        if averaging:
            ex = grid_average(eps0_ * eps_r, 'x').reshape(-1)
            ey = grid_average(eps0_ * eps_r, 'y').reshape(-1)
        else:
            ex = eps0_ * eps_r.reshape(-1)
            ey = eps0_ * eps_r.reshape(-1)
        Txi = sp.diags(1 / ex, 0, shape=(M, M), format='csr')
        Tyi = sp.diags(1 / ey, 0, shape=(M, M), format='csr')
        A = Dxf @ Txi @ Dxb + Dyf @ Tyi @ Dyb + omega ** 2 * mu0_ * sp.eye(M, format='csr')
    derivs = {'Dyb': Dyb, 'Dxb': Dxb, 'Dxf': Dxf, 'Dyf': Dyf}
    return A.tocsr(), derivs


# linalg.py:122-149 ---------------------------------------------------------
def solver_direct(A, b, solver='scipy'):
    """Direct solve with the zero-RHS shortcut (linalg.py:122-149).

    Only the reference's ``'scipy'`` branch (SuperLU ``spsolve``) exists here;
    Pardiso is not available. The shortcut returns real zeros, as the
    reference does.
    """
    b = np.asarray(b).astype(np.complex128).reshape(-1)
    if not b.any():
        return np.zeros(b.shape)
    if solver.lower() not in ('scipy', 'pardiso'):
        raise ValueError('Invalid solver choice: {}, options are pardiso or '
                         'scipy'.format(str(solver)))
    return spl.spsolve(A.tocsc(), b)


def solver_complex2real(A11, A12, b):
    """x with A11 x + A12 conj(x) = b, as the real system of order 2N
    [[Re A11 + Re A12, -Im A11 + Im A12], [Im A11 + Im A12, Re A11 - Re A12]]
    [Re x; Im x] = [Re b; Im b] (linalg.py:152-186). The reference solves it
    with Pardiso (mtype=11); its scipy branch names spsolve and zeros without
    importing them, so this restates the evident intent with SuperLU."""
    b = np.asarray(b).astype(np.complex128).reshape(-1)
    N = b.size
    if not b.any():
        return np.zeros(b.shape, dtype=np.complex128)
    A11, A12 = sp.csr_matrix(A11), sp.csr_matrix(A12)
    Areal = sp.vstack((sp.hstack((A11.real + A12.real, -A11.imag + A12.imag)),
                       sp.hstack((A11.imag + A12.imag, A11.real - A12.real)))).tocsc()
    x = spl.spsolve(Areal, np.concatenate((b.real, b.imag)))
    return x[:N] + 1j * x[N:]


def stretch_symmetrizer(omega, L0, Nx, Ny, NPML, dl):
    """D = diag(sxf_i * syf_j) with D @ A complex-symmetric for both pols.

    Follows from createDws' dxf == -dxb.T and the diagonal stretch matrices of
    pml.py:79-97; used for the transpose (adjoint) solve A^T = D A D^-1.
    """
    sxf, _, syf, _ = sfactor_vectors(omega, L0, Nx, Ny, NPML, dl)
    return (sxf[:, None] * syf[None, :]).reshape(-1)


def plane_wave_eigenvalue(pol, omega, L0, dl, eps, kx, ky):
    """Closed-form eigenvalue of the periodic, PML-free, uniform-eps operator.

    For x_{ij} = exp(i(kx*i + ky*j)*dl) the periodic differences give
    Dxf Dxb -> (2 cos(kx dl) - 2)/dl^2 (and likewise in y).
    """
    lap = (2 * np.cos(kx * dl) - 2) / dl ** 2 + (2 * np.cos(ky * dl) - 2) / dl ** 2
    if pol == 'Ez':
        return lap / (MU_0 * L0) + omega ** 2 * EPSILON_0 * L0 * eps
    return lap / (EPSILON_0 * L0 * eps) + omega ** 2 * MU_0 * L0


class OracleSimulation:
    """Restatement of ``Simulation`` on the CPU (simulation.py:15-258).

    Constructor, eps_r setter, solve_fields, the Kerr/Born iteration, the
    flux probe and modal sources follow the cited reference methods; the
    solve always uses SuperLU (the reference's ``solver='scipy'``).
    """

    def __init__(self, omega, eps_r, dl, NPML, pol, L0=DEFAULT_LENGTH_SCALE):
        # simulation.py:17-43
        self.L0 = L0
        self.omega = float(omega)
        self.dl = float(dl)
        self.NPML = [int(n) for n in NPML]
        self.pol = pol
        assert self.L0 > 0
        assert len(self.NPML) == 2
        assert self.NPML[0] >= 0 and self.NPML[1] >= 0
        assert self.pol in ['Ez', 'Hz']
        Nx, Ny = eps_r.shape
        self.Nx, self.Ny = Nx, Ny
        self.mu_r = np.ones((Nx, Ny))
        self.src = np.zeros((Nx, Ny))
        self.xrange = [0, float(Nx * self.dl)]
        self.yrange = [0, float(Ny * self.dl)]
        self.eps_r = eps_r
        self.modes = []
        self.nonlinearity = []
        self.nonlinearity_de = []
        self.eps_nl = np.zeros(eps_r.shape)

    @property
    def eps_r(self):
        return self._eps_r

    @eps_r.setter
    def eps_r(self, new_eps):
        # simulation.py:79-89
        self._eps_r = new_eps
        self.A, self.derivs = construct_A(self.omega, self.xrange, self.yrange,
                                          self._eps_r, self.NPML, self.pol, self.L0)
        self.fields = {f: None for f in ['Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz']}
        self.fields_nl = {f: None for f in ['Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz']}

    # simulation.py:112-178
    def solve_fields(self, include_nl=False, averaging=True, solver='scipy'):
        eps0_ = EPSILON_0 * self.L0
        mu0_ = MU_0 * self.L0
        b = self.src * 1j * self.omega
        if not include_nl:
            eps_tot = self.eps_r
            X = solver_direct(self.A, b, solver=solver)
        else:
            eps_tot = self.eps_r + self.eps_nl
            X = solver_direct(self.A + self.Anl, b, solver=solver)
        Nx, Ny = self.src.shape
        D = self.derivs
        if self.pol == 'Hz':
            if averaging:
                ex = grid_average(eps0_ * eps_tot, 'x').reshape(-1)
                ey = grid_average(eps0_ * eps_tot, 'y').reshape(-1)
            else:
                ex = (eps0_ * eps_tot).reshape(-1)
                ey = (eps0_ * eps_tot).reshape(-1)
            Ex = (1 / 1j / self.omega * (D['Dyb'] @ X) / ey).reshape(Nx, Ny)
            Ey = (-1 / 1j / self.omega * (D['Dxb'] @ X) / ex).reshape(Nx, Ny)
            Hz = X.reshape(Nx, Ny)
            if not include_nl:
                self.fields.update(Ex=Ex, Ey=Ey, Hz=Hz)
            return Ex, Ey, Hz
        Hx = (-1 / 1j / self.omega / mu0_ * (D['Dyb'] @ X)).reshape(Nx, Ny)
        Hy = (1 / 1j / self.omega / mu0_ * (D['Dxb'] @ X)).reshape(Nx, Ny)
        Ez = X.reshape(Nx, Ny)
        if not include_nl:
            self.fields.update(Hx=Hx, Hy=Hy, Ez=Ez)
        return Hx, Hy, Ez

    # simulation.py:70-73, nonlinearity.py:8-34
    def add_nl(self, chi, nl_region, eps_scale=False, eps_max=None):
        chi_i = chi / np.square(self.L0)
        if eps_scale:
            if eps_max is None:
                raise AssertionError("Must provide eps_max when eps_scale is True")
            f = lambda e, eps: 3 * chi_i * nl_region * np.square(np.abs(e)) * ((eps - 1) / (eps_max - 1))
            fde = lambda e, eps: 3 * chi_i * nl_region * np.conj(e) * ((eps - 1) / (eps_max - 1))
        else:
            f = lambda e, eps: 3 * chi_i * nl_region * np.square(np.abs(e))
            fde = lambda e, eps: 3 * chi_i * nl_region * np.conj(e)
        self.nonlinearity.append(f)
        self.nonlinearity_de.append(fde)

    # simulation.py:57-68 (eps_nl and dnl_de)
    def compute_nl(self, e):
        self.eps_nl = np.zeros(self.eps_r.shape)
        self.dnl_de = np.zeros(self.eps_r.shape, dtype=np.complex128)
        for f in self.nonlinearity:
            self.eps_nl = self.eps_nl + f(e, self.eps_r)
        for fde in self.nonlinearity_de:
            self.dnl_de = self.dnl_de + fde(e, self.eps_r)
        M = self.Nx * self.Ny
        self.Anl = sp.diags(self.omega ** 2 * EPSILON_0 * self.L0 * self.eps_nl.reshape(-1),
                            0, shape=(M, M), format='csr')

    # nonlinear_solvers.py:13-54 (Ez branch)
    def born_solve(self, Estart=None, conv_threshold=1e-10, max_num_iter=50):
        conv_array = np.zeros((max_num_iter, 1))
        if self.pol != 'Ez':
            raise ValueError('Invalid polarization: {}'.format(self.pol))
        if Estart is None:
            if self.fields['Ez'] is None:
                _, _, Ez = self.solve_fields()
            else:
                Ez = copy.deepcopy(self.fields['Ez'])
        else:
            Ez = Estart
        for istep in range(max_num_iter):
            Eprev = Ez
            self.compute_nl(Eprev)
            Hx, Hy, Ez = self.solve_fields(include_nl=True)
            convergence = np.linalg.norm(Ez - Eprev) / np.linalg.norm(Ez)
            conv_array[istep] = convergence
            if convergence < conv_threshold:
                break
        self.fields_nl.update(Hx=Hx, Hy=Hy, Ez=Ez)
        return Hx, Hy, Ez, conv_array

    # nonlinear_solvers.py:57-149 (Ez branch) with nl_eq_and_jac (:118-149):
    # J11 = A + Anl + diag(dAde E), J12 = diag(conj(dAde) E), dAde =
    # omega^2 eps0' dnl_de, and E <- E - Ediff with [J11 J12; conj] Ediff = f
    def newton_solve(self, Estart=None, conv_threshold=1e-10, max_num_iter=50):
        conv_array = np.zeros((max_num_iter, 1))
        if self.pol != 'Ez':
            raise ValueError('Invalid polarization: {}'.format(self.pol))
        if Estart is None:
            if self.fields['Ez'] is None:
                _, _, Ez = self.solve_fields()
            else:
                Ez = copy.deepcopy(self.fields['Ez'])
        else:
            Ez = Estart
        M = self.Nx * self.Ny
        for istep in range(max_num_iter):
            Eprev = Ez
            self.compute_nl(Eprev)
            Anl = self.A + self.Anl
            e = Eprev.reshape(-1)
            fE = Anl.dot(e) - self.src.reshape(-1) * 1j * self.omega
            dAde = self.dnl_de.reshape(-1) * self.omega ** 2 * EPSILON_0 * self.L0
            J11 = Anl + sp.diags(dAde * e, 0, shape=(M, M), format='csr')
            J12 = sp.diags(np.conj(dAde) * e, 0, shape=(M, M), format='csr')
            Ediff = solver_complex2real(J11, J12, fE)
            Ez = Eprev - Ediff.reshape(self.Nx, self.Ny)
            convergence = np.linalg.norm(Ez - Eprev) / np.linalg.norm(Ez)
            conv_array[istep] = convergence
            if convergence < conv_threshold:
                break
        self.compute_nl(Ez)
        Hx, Hy, Ez = self.solve_fields(include_nl=True)
        self.fields_nl.update(Hx=Hx, Hy=Hy, Ez=Ez)
        return Hx, Hy, Ez, conv_array

    # simulation.py:260-315
    def flux_probe(self, direction_normal, center, width, nl=False):
        if direction_normal == "x":
            ix = [center[0], center[0] + 1]
            iy = [int(center[1] - width / 2), int(center[1] + width / 2)]
        elif direction_normal == "y":
            ix = [int(center[0] - width / 2), int(center[0] + width / 2)]
            iy = [center[1], center[1] + 1]
        else:
            raise ValueError("The value of direction_normal is neither x nor y!")
        F = self.fields_nl if nl else self.fields
        win = (slice(ix[0], ix[1] + 1), slice(iy[0], iy[1] + 1))
        cut = (slice(ix[0], ix[1]), slice(iy[0], iy[1]))
        if self.pol == 'Ez':
            Ez_x = grid_average(F['Ez'][win], 'x')[:-1, :-1]
            Ez_y = grid_average(F['Ez'][win], 'y')[:-1, :-1]
            if direction_normal == "x":
                return self.dl * np.sum(-0.5 * np.real(Ez_x * np.conj(F['Hy'][cut])))
            return self.dl * np.sum(0.5 * np.real(Ez_y * np.conj(F['Hx'][cut])))
        Hz_x = grid_average(F['Hz'][win], 'x')[:-1, :-1]
        Hz_y = grid_average(F['Hz'][win], 'y')[:-1, :-1]
        if direction_normal == "x":
            return self.dl * np.sum(0.5 * np.real(F['Ey'][cut] * np.conj(Hz_x)))
        return self.dl * np.sum(-0.5 * np.real(F['Ex'][cut] * np.conj(Hz_y)))

    # simulation.py:45-55
    def add_mode(self, neff, direction_normal, center, width, scale=1, order=1):
        self.modes.append(OracleMode(neff, direction_normal, center, width, scale, order))

    def setup_modes(self):
        for m in self.modes:
            m.setup_src(self)


class OracleMode:
    """Modal line source (source/mode.py:9-107), ARPACK shift-invert eigs."""

    def __init__(self, neff, direction_normal, center, width, scale, order=1):
        self.neff = neff
        self.direction_normal = direction_normal
        self.center = center
        self.width = width
        self.scale = scale
        self.order = order

    def setup_src(self, sim):
        self.compute_normalization(sim)
        self.insert_mode(sim, sim.src)

    def compute_normalization(self, sim):
        # mode.py:26-60
        eps0 = sim.eps_r
        Nx, Ny = eps0.shape
        eps_max = np.max(np.abs(eps0))
        norm_eps = np.ones((Nx, Ny))
        sim_norm = copy.deepcopy(sim)
        new_center = list(self.center)
        if self.direction_normal == "x":
            norm_eps[:, eps0[self.center[0], :] > 1] = eps_max
            new_center[0] = Nx - new_center[0]
        elif self.direction_normal == "y":
            norm_eps[eps0[:, self.center[1]] > 1, :] = eps_max
            new_center[1] = Ny - new_center[1]
        else:
            raise ValueError("The value of direction_normal is not x or y!")
        sim_norm.eps_r = norm_eps
        self.insert_mode(sim_norm, sim_norm.src)
        sim_norm.solve_fields()
        sim.W_in = sim_norm.flux_probe(self.direction_normal, new_center, self.width)
        sim.E2_in = np.sum(np.square(np.abs(sim_norm.fields['Ez'])) * np.abs(sim_norm.src))

    def insert_mode(self, sim, destination):
        # mode.py:62-107
        eps0_ = EPSILON_0 * sim.L0
        mu0_ = MU_0 * sim.L0
        if self.direction_normal == "x":
            ix = [self.center[0], self.center[0] + 1]
            iy = [int(self.center[1] - self.width / 2), int(self.center[1] + self.width / 2)]
        else:
            ix = [int(self.center[0] - self.width / 2), int(self.center[0] + self.width / 2)]
            iy = [self.center[1], self.center[1] + 1]
        eps_r = sim.eps_r[ix[0]:ix[1], iy[0]:iy[1]]
        N = eps_r.size
        Dxb = createDws('x', 'b', [sim.dl], [N])
        Dxf = createDws('x', 'f', [sim.dl], [N])
        veps = eps0_ * eps_r.reshape(-1)
        veps_x = eps0_ * grid_average(eps_r, 'x').reshape(-1)
        Teps = sp.diags(veps, 0, shape=(N, N), format='csr')
        Tepsxinv = sp.diags(1 / veps_x, 0, shape=(N, N), format='csr')
        if sim.pol == 'Ez':
            A = np.square(sim.omega) * mu0_ * Teps + Dxf @ Dxb
        else:
            A = np.square(sim.omega) * mu0_ * Teps + Teps @ Dxf @ Tepsxinv @ Dxb
        beta = sim.omega * np.sqrt(mu0_ * eps0_) * self.neff
        vals, vecs = spl.eigs(A, k=self.order, sigma=np.square(beta), which='LM')
        src = vecs if self.order == 1 else vecs[:, self.order - 1]
        src = src * self.scale
        src = src.reshape((1, -1)) if self.direction_normal == 'x' else src.reshape((-1, 1))
        destination[ix[0]:ix[1], iy[0]:iy[1]] = np.abs(src) * np.sign(np.real(src))


def adjoint_gradient(sim, probe_mask):
    """dJ/deps for J = sum(|Ez|^2 * probe_mask), Ez polarisation.

    dA/deps_p = omega^2 eps0' e_p e_p^T, so dJ/deps_p = -2 Re(lam_p omega^2
    eps0' E_p) with A^T lam = conj(E) * mask. A^T is solved directly (CSR
    transpose), not through the symmetrizer, to stay independent of the
    device path.
    """
    _, _, Ez = sim.solve_fields()
    g = (np.conj(Ez) * probe_mask).reshape(-1)
    lam = spl.spsolve(sim.A.T.tocsc(), g)
    w2e = sim.omega ** 2 * EPSILON_0 * sim.L0
    grad = -2 * np.real(lam * w2e * Ez.reshape(-1))
    J = float(np.sum(np.abs(Ez) ** 2 * probe_mask))
    return J, grad.reshape(sim.Nx, sim.Ny)
