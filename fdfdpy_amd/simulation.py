"""Drop-in ``Simulation`` for fdfdpy's 2D solve path (fdfdpy/simulation.py).

Same constructor, ``eps_r`` property and ``solve_fields`` as the reference;
the operator lives on the GPU (assembled by a HIP kernel on every ``eps_r``
assignment) and ``solve_fields`` runs the device block-sweep direct solver.
Factors stay resident between calls until ``eps_r`` changes, which the
reference lists as a to-do (README.md:103).
"""

from __future__ import annotations

import time

import numpy as np
import torch

from . import _lib
from .constants import DEFAULT_LENGTH_SCALE, DEFAULT_SOLVER, EPSILON_0, MU_0
from .device import MAX_BLOCK, DeviceOperator
from .csr import operator_csr, stretched_derivatives
from .source import mode

FIELD_NAMES = ['Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz']
SOLVER_NAMES = ('hip', 'pardiso', 'scipy')
PRECISIONS = {'fp64': 'fp64', 'mixed': 'c64', 'c64': 'c64'}


class InputError(ValueError, AssertionError):
    """Invalid Simulation input.

    The reference asserts (simulation.py:249-258, AssertionError) while its
    tests expect ValueError (tests/test_simulation.py:29-81); this class is
    both.
    """


def _is_scalar(v):
    return np.isscalar(v) or (isinstance(v, np.ndarray) and v.ndim == 0) or \
        (isinstance(v, torch.Tensor) and v.dim() == 0)


class Simulation:
    """2D FDFD simulation (simulation.py:15-436, solve path).

    Parameters follow simulation.py:17: omega (rad/s), eps_r ((Nx, Ny)
    array), dl (cell size in units of L0), NPML ([Nx_pml, Ny_pml]), pol
    ('Ez' or 'Hz'), L0 (m). ``device`` selects the GPU (default: current).
    """

    def __init__(self, omega, eps_r, dl, NPML, pol, L0=DEFAULT_LENGTH_SCALE, device=None):
        self._check_raw_inputs(omega, eps_r, dl, NPML, pol, L0)
        self.L0 = L0
        self.omega = float(omega)
        self.dl = float(dl)
        self.NPML = [int(n) for n in NPML]
        self.pol = pol
        self._check_inputs()

        (Nx, Ny) = tuple(eps_r.shape)
        self.Nx = Nx
        self.Ny = Ny
        if 2 * self.NPML[0] > Nx or 2 * self.NPML[1] > Ny:
            raise InputError("NPML {} does not fit the {}x{} domain".format(self.NPML, Nx, Ny))
        if min(Nx, Ny) > MAX_BLOCK:
            raise InputError("grid {}x{}: the shorter side must be <= {} cells (the direct solver's "
                             "block size)".format(Nx, Ny, MAX_BLOCK))
        self.mu_r = np.ones((self.Nx, self.Ny))
        self.src = np.zeros((self.Nx, self.Ny))
        self.xrange = [0, float(Nx * self.dl)]
        self.yrange = [0, float(Ny * self.dl)]
        self._op = DeviceOperator((Nx, Ny), self.omega, self.dl, self.NPML, self.pol, self.L0, device)
        self.device = self._op.device

        # construct the system matrix (on the device)
        self.eps_r = eps_r

        self.modes = []
        self.nonlinearity = []
        self.eps_nl = np.zeros(eps_r.shape)
        self.dnl_de = np.zeros(eps_r.shape)
        self.dnl_deps = np.zeros(eps_r.shape)
        self.last_report = None

    # ------------------------------------------------------------------ checks
    @staticmethod
    def _check_raw_inputs(omega, eps_r, dl, NPML, pol, L0):
        """Type and sign checks the reference test suite asks for."""
        if not _is_scalar(omega) or not np.isreal(omega) or not float(omega) > 0:
            raise InputError("omega must be a positive number, was supplied {}".format(omega))
        if not _is_scalar(dl) or not np.isreal(dl) or not float(dl) > 0:
            raise InputError("dl must be a positive number, was supplied {}".format(dl))
        if not isinstance(eps_r, (np.ndarray, torch.Tensor)) or len(eps_r.shape) != 2:
            raise InputError("eps_r must be a 2D numpy array or tensor")
        if _is_scalar(NPML) or len(NPML) != 2:
            raise InputError("NPML must be a list of length 2, was supplied {}".format(NPML))
        if not isinstance(pol, str):
            raise InputError("pol must be one of 'Ez' or 'Hz'")
        if not _is_scalar(L0) or not float(L0) > 0:
            raise InputError("L0 must be a positive number, was supplied {},".format(L0))
        Simulation._check_eps(eps_r)

    def _check_inputs(self):
        # simulation.py:249-258
        if not self.L0 > 0:
            raise InputError("L0 must be a positive number, was supplied {},".format(str(self.L0)))
        if len(self.NPML) != 2:
            raise InputError("NPML must be a list of length 2")
        if not (self.NPML[0] >= 0 and self.NPML[1] >= 0):
            raise InputError("both elements of NPML must be >= 0")
        if self.pol not in ['Ez', 'Hz']:
            raise InputError("pol must be one of 'Ez' or 'Hz'")

    @staticmethod
    def _check_eps(eps):
        e = eps.detach().cpu().numpy() if isinstance(eps, torch.Tensor) else np.asarray(eps)
        if not np.all(np.isfinite(e)):
            raise InputError("eps_r must be finite")
        if not np.iscomplexobj(e) and np.any(e <= 0):
            raise InputError("real eps_r must be positive (use a complex eps_r for lossy or metallic media)")

    # ------------------------------------------------------------------ eps_r
    @property
    def eps_r(self):
        return self.__eps_r

    @eps_r.setter
    def eps_r(self, new_eps):
        # simulation.py:79-89: assigning eps_r rebuilds the operator
        if tuple(new_eps.shape) != (self.Nx, self.Ny):
            raise InputError("eps_r has shape {}, expected {}".format(tuple(new_eps.shape), (self.Nx, self.Ny)))
        self._check_eps(new_eps)
        self.__eps_r = new_eps
        self._op.set_eps(new_eps)
        self.fields = {f: None for f in FIELD_NAMES}
        self.fields_nl = {f: None for f in FIELD_NAMES}

    def reset_eps(self, new_eps):
        """simulation.py:91-101."""
        self.eps_r = new_eps

    @property
    def operator(self):
        """The device operator (band storage + resident factors)."""
        return self._op

    @property
    def A(self):
        """The system matrix as a host scipy CSR matrix, as fdfdpy keeps it
        (simulation.py:82-86): copied from the device operator on first use
        after each eps_r assignment (the linear operator; Kerr terms are
        separate, as there)."""
        key = (id(self._op), self._op.linear_version)
        if getattr(self, '_A_key', None) != key:
            self._A_csr, self._A_key = operator_csr(self._op), key
        return self._A_csr

    @property
    def derivs(self):
        """{'Dxf', 'Dxb', 'Dyf', 'Dyb'}: the stretched difference operators
        of construct_A (simulation.py:87, linalg.py:55-58) as host CSR
        matrices, built on first use."""
        if getattr(self, '_derivs', None) is None:
            self._derivs = stretched_derivatives(self.omega, (self.Nx, self.Ny), self.NPML, self.dl, self.L0)
        return self._derivs

    # ------------------------------------------------------------------ solve
    def _src_tensor(self):
        s = self.src
        t = s if isinstance(s, torch.Tensor) else torch.from_numpy(np.asarray(s))
        return t.to(device=self.device, dtype=torch.complex128)

    def _out(self, t):
        return t if isinstance(self.src, torch.Tensor) else t.cpu().numpy()

    def _solve_primary(self, precision, tol, max_iter, timing):
        b = self._src_tensor() * (1j * self.omega)
        if not bool(torch.any(b != 0)):
            # linalg.py:128-129 (the reference returns real zeros; complex here)
            return torch.zeros_like(b)
        t0 = time.time()
        x, rep = self._op.solve(b, precision=precision, tol=tol, max_iter=max_iter)
        self.last_report = rep
        if timing:
            print('Linear system solve took {:.2f} seconds'.format(time.time() - t0))
        return x

    def solve_fields(self, include_nl=False, timing=False, averaging=True, solver=DEFAULT_SOLVER,
                     matrix_format='csr', precision='fp64', tol=1e-12, max_iter=60):
        """Solve A x = i omega src (simulation.py:112-178).

        Returns (Hx, Hy, Ez) for pol='Ez' or (Ex, Ey, Hz) for pol='Hz', numpy
        complex128 arrays when ``src`` is a numpy array and device tensors when
        it is a tensor. ``solver`` accepts 'hip' and the reference's names
        'pardiso' and 'scipy', which all run the device direct solver;
        ``precision`` 'fp64' factors in FP64, 'mixed' in complex64 with FP64
        refinement (and an FP64 fallback). ``tol`` is the relative residual.
        """
        if str(solver).lower() not in SOLVER_NAMES:
            raise ValueError('Invalid solver choice: {}, options are hip, pardiso or scipy'.format(str(solver)))
        if precision not in PRECISIONS:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if self.pol not in ('Ez', 'Hz'):
            raise ValueError('Invalid polarization: {}'.format(str(self.pol)))
        op = self._op
        if include_nl:
            # A + Anl (simulation.py:121-124): the diagonal of the device
            # operator is set from sim.eps_nl, whatever it held before
            eps_tot = self._eps_total()
            op.set_kerr_diagonal(torch.as_tensor(np.real(np.asarray(self.eps_nl)), dtype=torch.float64))
        else:
            op.restore_linear()
            eps_tot = None
        X = self._solve_primary(PRECISIONS[precision], tol, max_iter, timing)
        f1, f2 = op.fields(X, eps_total=eps_tot, averaging=averaging)
        if self.pol == 'Hz':
            Ex, Ey, Hz = self._out(f1), self._out(f2), self._out(X)
            if not include_nl:
                self.fields['Ex'] = Ex
                self.fields['Ey'] = Ey
                self.fields['Hz'] = Hz
            return (Ex, Ey, Hz)
        Hx, Hy, Ez = self._out(f1), self._out(f2), self._out(X)
        if not include_nl:
            self.fields['Hx'] = Hx
            self.fields['Hy'] = Hy
            self.fields['Ez'] = Ez
        return (Hx, Hy, Ez)

    def _eps_total(self):
        eps = torch.as_tensor(self.eps_r, device=self.device).to(torch.complex128)
        enl = torch.as_tensor(self.eps_nl, device=self.device).to(torch.complex128)
        return (eps + enl).contiguous()

    # -------------------------------------------------------------- nonlinear
    def add_nl(self, chi, nl_region, nl_type='kerr', eps_scale=False, eps_max=None):
        """simulation.py:70-73 with nonlinearity.py:8-37 (Kerr only)."""
        from .nonlinear import Nonlinearity
        self.nonlinearity.append(Nonlinearity(chi / np.square(self.L0), nl_region, nl_type,
                                              eps_scale, eps_max))

    def compute_nl(self, e):
        """eps_nl for field e and the nonlinear diagonal on the device (simulation.py:57-68)."""
        from .nonlinear import compute_nl
        return compute_nl(self, e)

    def solve_fields_nl(self, timing=False, averaging=True, Estart=None, solver_nl='newton',
                        conv_threshold=1e-10, max_num_iter=50, solver=DEFAULT_SOLVER,
                        matrix_format='csr', precision='fp64'):
        """Nonlinear solve (simulation.py:180-247) on the device: Newton's
        method (the reference's default) or the Born iteration."""
        from .nonlinear import born_solve, newton_solve
        if self.pol != 'Ez':
            if self.pol == 'Hz':
                raise ValueError('the Kerr nonlinearity is supported for pol=Ez only '
                                 '(nonlinear_solvers.py:54 references an undefined name for Hz)')
            raise ValueError('Invalid polarization: {}'.format(str(self.pol)))
        if solver_nl == 'born':
            Hx, Hy, Ez, conv_array = born_solve(self, Estart, conv_threshold, max_num_iter,
                                                averaging=averaging, precision=PRECISIONS[precision])
        elif solver_nl == 'newton':
            Hx, Hy, Ez, conv_array = newton_solve(self, Estart, conv_threshold, max_num_iter,
                                                  averaging=averaging, precision=PRECISIONS[precision])
        elif solver_nl == 'LM':
            # simulation.py:199 calls an LM_solve that the reference never defines
            raise NotImplementedError("solver_nl='LM' is not defined by the reference either")
        else:
            raise AssertionError("solver must be one of {'born', 'newton', 'LM'}")
        self.fields_nl['Hx'] = Hx
        self.fields_nl['Hy'] = Hy
        self.fields_nl['Ez'] = Ez
        return (Hx, Hy, Ez, conv_array)

    def compute_index_shift(self):
        """simulation.py:103-110."""
        _ = self.solve_fields()
        _ = self.solve_fields_nl()
        index_nl = np.sqrt(np.real(np.asarray(self.eps_r) + np.asarray(self.eps_nl)))
        index_lin = np.sqrt(np.real(np.asarray(self.eps_r)))
        return np.abs(index_nl - index_lin)

    # ------------------------------------------------------------------ modes
    def add_mode(self, neff, direction_normal, center, width, scale=1, order=1):
        """simulation.py:50-55: a modal source definition, set up by setup_modes."""
        self.modes.append(mode(neff, direction_normal, center, width, scale=scale, order=order))

    def setup_modes(self):
        """simulation.py:45-48: normalization run (device solve) and mode
        profile into src, for every mode added."""
        for m in self.modes:
            m.setup_src(self)

    # ------------------------------------------------------------------ flux
    def flux_probe(self, direction_normal, center, width, nl=False):
        """Poynting flux through a line (simulation.py:260-315), on host arrays."""
        if direction_normal == "x":
            ix = [center[0], center[0] + 1]
            iy = [int(center[1] - width / 2), int(center[1] + width / 2)]
        elif direction_normal == "y":
            ix = [int(center[0] - width / 2), int(center[0] + width / 2)]
            iy = [center[1], center[1] + 1]
        else:
            raise ValueError("The value of direction_normal is neither x nor y!")
        F = self.fields_nl if nl else self.fields
        flux = self._flux_on_device(F, direction_normal, ix, iy)
        if flux is not None:
            return flux
        get = lambda k: F[k].cpu().numpy() if isinstance(F[k], torch.Tensor) else F[k]
        win = (slice(ix[0], ix[1] + 1), slice(iy[0], iy[1] + 1))
        cut = (slice(ix[0], ix[1]), slice(iy[0], iy[1]))

        def avg(a, axis):
            return ((np.roll(a, 1, axis=axis) + a) / 2)[:-1, :-1]

        if self.pol == 'Ez':
            Ez = get('Ez')[win]
            if direction_normal == "x":
                return self.dl * np.sum(-0.5 * np.real(avg(Ez, 0) * np.conj(get('Hy')[cut])))
            return self.dl * np.sum(0.5 * np.real(avg(Ez, 1) * np.conj(get('Hx')[cut])))
        Hz = get('Hz')[win]
        if direction_normal == "x":
            return self.dl * np.sum(0.5 * np.real(get('Ey')[cut] * np.conj(avg(Hz, 0))))
        return self.dl * np.sum(-0.5 * np.real(get('Ex')[cut] * np.conj(avg(Hz, 1))))


    def _flux_on_device(self, F, direction_normal, ix, iy):
        """The flux sum by the library (fdfd_flux) when the fields it needs
        are device tensors and numpy's slicing of the probe window would not
        be clipped by the grid edge; None otherwise (host arrays below)."""
        kind = {('Ez', 'x'): 0, ('Ez', 'y'): 1, ('Hz', 'x'): 2, ('Hz', 'y'): 3}[(self.pol, direction_normal)]
        e_name, h_name = [('Ez', 'Hy'), ('Ez', 'Hx'), ('Ey', 'Hz'), ('Ex', 'Hz')][kind]
        e, h = F[e_name], F[h_name]
        if not (isinstance(e, torch.Tensor) and isinstance(h, torch.Tensor) and e.is_cuda and h.is_cuda):
            return None
        if direction_normal == 'x':
            i0, j0, n = ix[0], iy[0], iy[1] - iy[0]
            inside = i0 >= 0 and i0 + 2 <= self.Nx and j0 >= 0 and iy[1] + 1 <= self.Ny
        else:
            i0, j0, n = ix[0], iy[0], ix[1] - ix[0]
            inside = j0 >= 0 and j0 + 2 <= self.Ny and i0 >= 0 and ix[1] + 1 <= self.Nx
        if not inside or n < 0:
            return None
        e = e.to(torch.complex128).contiguous()
        h = h.to(torch.complex128).contiguous()
        out = torch.empty(1, dtype=torch.float64, device=e.device)
        lib = _lib.lib()
        _lib.check(lib.fdfd_flux(self.Nx, self.Ny, self.dl, kind, e.data_ptr(), h.data_ptr(), int(i0), int(j0),
                                 int(n), out.data_ptr(), torch.cuda.current_stream(e.device).cuda_stream))
        return float(out.item())


__all__ = ['Simulation', 'InputError', 'EPSILON_0', 'MU_0']
