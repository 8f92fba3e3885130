"""Device-resident FDFD operator with its factors (host side of the C ABI).

One ``DeviceOperator`` holds, in HBM, what fdfdpy keeps as ``Simulation.A``
(linalg.py:41-107) -- here the 5-plane band storage written by the assembly
kernel -- plus the factors that fdfdpy rebuilds and frees on every
``solver_direct`` call (linalg.py:135-138). Factors stay resident until the
operator changes, so further sources, adjoint solves and Born re-solves reuse
them.

Orientation: the block sweep runs along one axis with blocks along the
other. The operator of the transposed grid is the same physics with x and y
exchanged (construct_A treats both axes alike), so the grid is transposed on
the device when that gives a damped sweep axis (NPML >= 1) or a smaller block.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import torch

from . import _lib as L

MAX_BLOCK = 2048   # the library's largest block (the shorter grid side)


def _vp(t):
    """Raw device pointer of a tensor handed to the C ABI.

    The ABI reads memory as it lies, so lazily conjugated/negated views and
    strided views are refused here rather than read wrongly.
    """
    if t is None:
        return ctypes.c_void_p(0)
    if t.is_conj() or t.is_neg() or not t.is_contiguous():
        raise ValueError("tensor passed to the C ABI must be contiguous with no conj/neg view bits")
    return ctypes.c_void_p(t.data_ptr())


def _dense(t):
    return t.resolve_conj().resolve_neg().contiguous()


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def choose_transpose(Nx, Ny, NPML):
    """True when the sweep should run along y (blocks of size Nx)."""
    def score(n_block, npml_sweep):
        if n_block > MAX_BLOCK:
            return None
        return (0 if npml_sweep >= 1 else 1, n_block)
    a = score(Ny, NPML[0])      # sweep along x, blocks of Ny
    b = score(Nx, NPML[1])      # sweep along y, blocks of Nx
    if a is None and b is None:
        raise ValueError(f"grid {Nx}x{Ny}: the shorter side must be <= {MAX_BLOCK} cells")
    if a is None:
        return True
    if b is None:
        return False
    return b < a


@dataclass
class SolveReport:
    precision: str
    iterations: int
    converged: bool
    max_rel_residual: float
    rel_residual: list
    fell_back: bool = False     # solved with a factor other than the requested one
    pivoted: bool = False       # the factor pivots inside its 32x32 pivot tiles
    restarts: int = 0           # persistent sweeps timed out; redone per block
    sweeps: int = 0             # block sweeps T^-1 r of the solve that produced x
    failed_factors: list = field(default_factory=list)   # (precision, pivoted) with a zero/non-finite pivot
    factor_seconds: float = 0.0
    solve_seconds: float = 0.0
    history: list = field(default_factory=list)


class FactorEngine:
    """Step-level batching of factorizations (fdfd_engine_* in the C ABI).

    Factorizations submitted from any thread share the engine's Gauss-Jordan
    step launches; wait() blocks until one is complete.
    """

    def __init__(self, device):
        self.lib = L.lib()
        self.device = torch.device(device)
        idx = self.device.index if self.device.index is not None else torch.cuda.current_device()
        h = self.lib.fdfd_engine_create(idx)
        if not h:
            raise L.FdfdError(L.FDFD_ERR_HIP, self.lib.fdfd_last_error().decode())
        self.handle = ctypes.c_void_p(h)

    def submit(self, grid, op, fac, precision, work):
        """precision: an FDFD_PREC_* code, optionally | FACTOR_PIVOT."""
        t = ctypes.c_uint64(0)
        L.check(self.lib.fdfd_engine_submit(self.handle, ctypes.byref(grid), _vp(op), _vp(fac),
                                            int(precision), _vp(work), ctypes.byref(t)))
        return t.value

    def wait(self, ticket):
        L.check(self.lib.fdfd_engine_wait(self.handle, ctypes.c_uint64(ticket)))

    @property
    def launches(self):
        return int(self.lib.fdfd_engine_launches(self.handle))

    def close(self):
        if self.handle:
            self.lib.fdfd_engine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceOperator:
    """Assembled operator + resident factors for one (grid, omega, eps)."""

    def __init__(self, shape, omega, dl, NPML, pol, L0, device=None):
        Nx, Ny = int(shape[0]), int(shape[1])
        if pol not in L.POL:
            raise ValueError("pol must be one of 'Ez' or 'Hz'")
        self.shape = (Nx, Ny)
        self.pol = pol
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if self.device.type != "cuda":
            raise ValueError("DeviceOperator needs a GPU device")
        self.transposed = choose_transpose(Nx, Ny, NPML)
        nx, ny = (Ny, Nx) if self.transposed else (Nx, Ny)
        npx, npy = (NPML[1], NPML[0]) if self.transposed else (NPML[0], NPML[1])
        self.grid = L.FdfdGrid(nx, ny, int(npx), int(npy), L.POL[pol], 0, float(omega), float(dl), float(L0))
        # grid of the caller's orientation (secondary fields)
        self.grid_user = L.FdfdGrid(Nx, Ny, int(NPML[0]), int(NPML[1]), L.POL[pol], 0,
                                    float(omega), float(dl), float(L0))
        self.lib = L.lib()
        nbytes = self.lib.fdfd_op_bytes(ctypes.byref(self.grid))
        if nbytes == 0:
            raise ValueError(self.lib.fdfd_last_error().decode())
        self.op = torch.empty(nbytes // 8, dtype=torch.float64, device=self.device)
        self.eps = None          # complex128, solver orientation
        self.eps_user = None     # complex128, caller orientation
        self.base_c = None       # linear diagonal plane (for Kerr updates)
        self.version = 0         # bumped whenever the operator's diagonal changes
        self.factors = {}        # precision -> (uint8 tensor, operator version, pivoted)
        self.factor_seconds = {}
        self.engine = None       # FactorEngine shared with other operators, if any

    # -- sizes ----------------------------------------------------------
    def factor_bytes(self, precision):
        return int(self.lib.fdfd_factor_bytes(ctypes.byref(self.grid), L.PREC[precision]))

    def op_bytes(self):
        return self.op.numel() * 8

    # -- orientation helpers ---------------------------------------------
    def to_solver(self, v):
        """(..., Nx, Ny) caller layout -> contiguous solver layout."""
        return _dense(v.transpose(-1, -2) if self.transposed else v)

    def to_user(self, v):
        return _dense(v.transpose(-1, -2) if self.transposed else v)

    # -- assembly ----------------------------------------------------------
    def set_eps(self, eps):
        """Assemble A for a new permittivity (the eps_r setter, simulation.py:79-89)."""
        e = torch.as_tensor(eps, device=self.device)
        if e.shape != self.shape:
            raise ValueError(f"eps_r has shape {tuple(e.shape)}, expected {self.shape}")
        e = e.to(torch.complex128)
        self.eps_user = _dense(e)
        self.eps = self.to_solver(e)
        L.check(self.lib.fdfd_assemble(ctypes.byref(self.grid), _vp(self.eps), _vp(self.op),
                                       _stream(self.device)))
        M = self.shape[0] * self.shape[1]
        self.base_c = self.op[: 2 * M].clone()
        self.version += 1
        self.linear_version = self.version
        self.factors.clear()
        self.factor_seconds.clear()

    def restore_linear(self):
        """Undo Kerr diagonal updates (the linear operator is kept, as Simulation.A is)."""
        if self.version != self.linear_version:
            M = self.shape[0] * self.shape[1]
            self.op[: 2 * M].copy_(self.base_c)
            self.version = self.linear_version

    # -- factor --------------------------------------------------------------
    def factor(self, precision="fp64", refresh=False, pivot=False, stepwise=False):
        """Resident factor for this precision; built for the current operator
        when absent, or when refresh is set and the resident one belongs to
        an older operator or another pivoting mode. pivot=True pivots inside
        each 32x32 pivot tile (the fallback when the diagonal-pivot factor
        does not refine). stepwise=True forces one launch per Gauss-Jordan
        step instead of the persistent launch."""
        if precision in self.factors:
            fac, ver, piv = self.factors[precision]
            if not refresh or (ver == self.version and piv == pivot):
                return fac
        g = ctypes.byref(self.grid)
        p = L.PREC[precision] | (L.FACTOR_PIVOT if pivot else 0) | (L.FACTOR_STEPWISE if stepwise else 0)
        fac = torch.empty(self.lib.fdfd_factor_bytes(g, L.PREC[precision]), dtype=torch.uint8,
                          device=self.device)
        work = torch.empty(self.lib.fdfd_factor_workspace_bytes(g, L.PREC[precision]), dtype=torch.uint8,
                           device=self.device)
        s = _stream(self.device)
        t0 = time.perf_counter()
        if self.engine is not None:
            torch.cuda.current_stream(self.device).synchronize()   # operator complete
            self.engine.wait(self.engine.submit(self.grid, self.op, fac, p, work))
            L.check(self.lib.fdfd_factor_wrap(g, _vp(self.op), _vp(fac), p, _vp(work), s))
            L.check(self.lib.fdfd_factor_status(g, p, _vp(work), s))
        else:
            L.check(self.lib.fdfd_factor(g, _vp(self.op), _vp(fac), p, _vp(work), s))
            L.check(self.lib.fdfd_factor_status(g, p, _vp(work), s))
        self.factor_seconds[precision] = time.perf_counter() - t0
        self.factors[precision] = (fac, self.version, pivot)
        return fac

    def drop_factor(self, precision=None):
        if precision is None:
            self.factors.clear()
        else:
            self.factors.pop(precision, None)

    # -- solve ---------------------------------------------------------------
    def solve(self, b, precision="fp64", tol=1e-12, max_iter=60, transpose=False,
              fallback=True):
        """x with A x = b (or A^T x = b), b: (nrhs, Nx, Ny) or (Nx, Ny) complex.

        precision 'c64' factors in complex64 and refines in FP64. When the
        refinement stalls the solve falls back to the FP64 factor, and then
        to the FP64 factor with pivoting inside the pivot tiles.
        """
        squeeze = b.dim() == 2
        bb = b.reshape((1,) + self.shape) if squeeze else b
        nrhs = bb.shape[0]
        bs = self.to_solver(bb.to(device=self.device, dtype=torch.complex128))
        x = torch.empty_like(bs)
        # attempts in order: (precision, pivoted, refactor first)
        plan = [(precision, False, False)]
        if precision in self.factors and self.factors[precision][1] != self.version:
            # the resident factor belongs to an earlier diagonal (Born steps):
            # it preconditions the refinement first, then is rebuilt
            plan.append((precision, False, True))
        if precision == "c64" and fallback:
            plan.append(("fp64", False, True))
        if fallback:
            # last resort: partial pivoting inside the 32x32 pivot tiles
            plan.append(("fp64", True, True))
        rep, failed, spent = None, [], [0.0, 0.0, 0]
        for prec, piv, refresh in plan:
            if rep is not None and rep.converged:
                break
            try:
                if refresh:
                    self.factor(prec, refresh=True, pivot=piv)
                rep2 = self._solve_one(bs, x, nrhs, prec, tol, max_iter, transpose)
            except L.FdfdError as e:
                if e.code != L.FDFD_ERR_NUMERIC:
                    raise
                # a zero or non-finite pivot: this factor is unusable
                self.drop_factor(prec)
                failed.append((prec, piv))
                continue
            rep2.fell_back = prec != precision or piv
            rep2.pivoted = self.factors[prec][2]
            spent = [spent[0] + rep2.factor_seconds, spent[1] + rep2.solve_seconds, spent[2] + rep2.restarts]
            rep = rep2
        if rep is None:
            raise L.FdfdError(L.FDFD_ERR_NUMERIC, f"every factor in the fallback order had a zero or "
                              f"non-finite pivot: {failed}")
        rep.factor_seconds, rep.solve_seconds, rep.restarts = spent
        rep.failed_factors = failed
        if not rep.converged:
            raise L.NotConverged(L.FDFD_ERR_NOT_CONVERGED,
                                 f"solve did not reach rel. residual {tol:g}: {rep.max_rel_residual:.3e} "
                                 f"after {rep.iterations} refinement steps ({rep.precision})")
        out = self.to_user(x)
        return (out[0] if squeeze else out), rep

    def _solve_one(self, bs, x, nrhs, precision, tol, max_iter, transpose):
        t0 = time.perf_counter()
        fac = self.factor(precision)
        tf = time.perf_counter() - t0 if precision not in self.factor_seconds else self.factor_seconds[precision]
        g = ctypes.byref(self.grid)
        p = L.PREC[precision]
        work = torch.empty(self.lib.fdfd_solve_workspace_bytes(g, p, nrhs), dtype=torch.uint8,
                           device=self.device)
        st = L.FdfdSolveStats()
        rel = (ctypes.c_double * nrhs)()
        t1 = time.perf_counter()
        rc = self.lib.fdfd_solve(g, _vp(self.op), _vp(fac), p, _vp(bs), _vp(x), nrhs, int(bool(transpose)),
                                 float(tol), int(max_iter), _vp(work), ctypes.byref(st), rel,
                                 _stream(self.device))
        if rc not in (L.FDFD_OK, L.FDFD_ERR_NOT_CONVERGED):
            L.check(rc)
        return SolveReport(precision=precision, iterations=st.iterations, converged=bool(st.converged),
                           max_rel_residual=st.max_rel_residual, rel_residual=list(rel), restarts=st.restarts,
                           sweeps=st.sweeps, factor_seconds=tf, solve_seconds=time.perf_counter() - t1)

    # -- stencil -------------------------------------------------------------
    def apply(self, x):
        """A x for (nrhs, Nx, Ny) or (Nx, Ny) complex x (caller layout)."""
        squeeze = x.dim() == 2
        xx = x.reshape((1,) + self.shape) if squeeze else x
        xs = self.to_solver(xx.to(device=self.device, dtype=torch.complex128))
        y = torch.empty_like(xs)
        L.check(self.lib.fdfd_apply(ctypes.byref(self.grid), _vp(self.op), _vp(xs), _vp(y), xs.shape[0],
                                    _stream(self.device)))
        out = self.to_user(y)
        return out[0] if squeeze else out

    def fields(self, x, eps_total=None, averaging=True):
        """Secondary fields (simulation.py:133-175) of primary fields x (caller layout)."""
        squeeze = x.dim() == 2
        xx = _dense((x.reshape((1,) + self.shape) if squeeze else x).to(torch.complex128))
        f1 = torch.empty_like(xx)
        f2 = torch.empty_like(xx)
        e = self.eps_user if eps_total is None else _dense(eps_total.to(torch.complex128))
        L.check(self.lib.fdfd_fields(ctypes.byref(self.grid_user), _vp(e), _vp(xx), _vp(f1), _vp(f2),
                                     xx.shape[0], int(bool(averaging)), _stream(self.device)))
        if squeeze:
            return f1[0], f2[0]
        return f1, f2

    def kerr_update(self, e, chi_region):
        """op.C = linear C + omega^2 eps0 L0 chi_region |e|^2; returns eps_nl (caller layout)."""
        self.version += 1
        es = self.to_solver(e.to(torch.complex128))
        chi = self.to_solver(chi_region.to(device=self.device, dtype=torch.float64))
        enl = torch.empty(chi.shape, dtype=torch.float64, device=self.device)
        L.check(self.lib.fdfd_kerr_update(ctypes.byref(self.grid), _vp(self.op), _vp(self.base_c), _vp(es),
                                          _vp(chi), _vp(enl), _stream(self.device)))
        return self.to_user(enl)

    def set_kerr_diagonal(self, eps_nl):
        """op.C = linear C + omega^2 eps0 L0 eps_nl for a given real eps_nl
        (caller layout): A + Anl of simulation.py:121-124."""
        e = torch.ones(self.shape, dtype=torch.complex128, device=self.device)
        return self.kerr_update(e, eps_nl)

    def grad_eps(self, lam, e):
        """-Re(lam * omega^2 eps0 L0 * e) (Ez), both in caller layout."""
        lam = _dense(lam.to(torch.complex128))
        e = _dense(e.to(torch.complex128))
        out = torch.empty(self.shape, dtype=torch.float64, device=self.device)
        L.check(self.lib.fdfd_grad_eps(ctypes.byref(self.grid_user), _vp(lam), _vp(e), _vp(out),
                                       _stream(self.device)))
        return out
