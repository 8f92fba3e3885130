"""GPU parity of the HIP path against the oracle and the golden fixtures."""

import ctypes
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from conftest import GOLDEN
from fdfdpy_amd import _lib as L
from fdfdpy_amd import configs as C
from fdfdpy_amd.device import DeviceOperator
from fdfdpy_amd.simulation import Simulation
from oracle import EPSILON_0, OracleSimulation, adjoint_gradient, solver_direct

pytestmark = pytest.mark.gpu
OMEGA = C.omega_of(1.55e-6)
DEV = "cuda:0"


def relerr(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("mode", [0, 1])
def test_mfma_lane_layout(mode):
    rng = np.random.default_rng(mode)
    a = torch.tensor(rng.standard_normal((16, 4)), device=DEV)
    b = torch.tensor(rng.standard_normal((4, 16)), device=DEV)
    d = torch.zeros((16, 16), dtype=torch.float64, device=DEV)
    lib = L.lib()
    L.check(lib.fdfd_selftest_mfma(mode, ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                   ctypes.c_void_p(d.data_ptr()), None))
    torch.cuda.synchronize()
    tol = 1e-13 if mode == 0 else 1e-5
    assert np.abs(d.cpu().numpy() - (a @ b).cpu().numpy()).max() < tol


@pytest.mark.parametrize("pol,N,npml", [
    ("Ez", (40, 37), [8, 6]), ("Hz", (36, 45), [7, 9]), ("Ez", (30, 50), [0, 6]),
    ("Hz", (64, 64), [10, 10]), ("Ez", (33, 17), [2, 3]),
])
def test_assembly_matches_oracle_operator(pol, N, npml):
    rng = np.random.default_rng(1)
    eps = 1 + 11.25 * (rng.random(N) < 0.5)
    sim = OracleSimulation(OMEGA, eps, 0.02, npml, pol)
    op = DeviceOperator(N, OMEGA, 0.02, npml, pol, 1e-6, DEV)
    op.set_eps(eps)
    x = rng.standard_normal((3,) + N) + 1j * rng.standard_normal((3,) + N)
    y = op.apply(torch.tensor(x, device=DEV)).cpu().numpy()
    for q in range(3):
        yref = (sim.A @ x[q].reshape(-1)).reshape(N)
        assert relerr(y[q], yref) < 1e-14


@pytest.mark.parametrize("pol,N,npml", [
    ("Ez", (40, 37), [8, 6]), ("Hz", (36, 45), [7, 9]), ("Ez", (30, 50), [0, 6]), ("Ez", (70, 64), [10, 10]),
])
@pytest.mark.parametrize("precision", ["fp64", "c64"])
def test_solve_matches_oracle(pol, N, npml, precision):
    rng = np.random.default_rng(2)
    eps = 1 + 11.25 * (rng.random(N) < 0.5)
    sim = OracleSimulation(OMEGA, eps, 0.02, npml, pol)
    op = DeviceOperator(N, OMEGA, 0.02, npml, pol, 1e-6, DEV)
    op.set_eps(eps)
    src = np.zeros((2,) + N)
    src[0, N[0] // 2, N[1] // 3] = 1
    src[1] = rng.standard_normal(N)
    b = src * 1j * OMEGA
    x, rep = op.solve(torch.tensor(b, device=DEV), precision=precision, tol=1e-13)
    assert rep.converged and rep.max_rel_residual <= 1e-13
    for q in range(2):
        xref = solver_direct(sim.A, b[q]).reshape(N)
        assert relerr(x[q].cpu().numpy(), xref) < 1e-10


@pytest.mark.parametrize("npml,precision,max_sweeps", [
    ([15, 15], "fp64", 3), ([15, 15], "c64", 5), ([0, 12], "fp64", 2), ([0, 12], "c64", 8),
    ([12, 0], "fp64", 2),
])
def test_wrap_correction_refinement_steps(npml, precision, max_sweeps):
    """Block sweeps a solve to 1e-10 needs with the x-wrap correction. The
    truncated correction (PML on both axes) is lagged, one sweep per step: a
    FP64 solve takes three (uncorrected, corrected, refined) and a complex64
    solve at most five. The exact correction (a periodic axis) takes two
    sweeps per step: one step in FP64, at most four in complex64. Without the
    correction the wrap costs tens of sweeps."""
    p = C.c3_request(3, n=128)
    N = p.shape
    op = DeviceOperator(N, p.omega, p.dl, npml, "Ez", p.L0, DEV)
    op.set_eps(p.eps)
    b = p.rhs()
    x, rep = op.solve(torch.tensor(b, device=DEV), precision=precision, tol=1e-10)
    print(f"npml {npml} {precision}: {rep.iterations} steps, {rep.sweeps} sweeps, residual {rep.max_rel_residual:.2e}")
    assert rep.converged and not rep.fell_back
    assert rep.sweeps <= max_sweeps, rep.sweeps
    assert rep.sweeps == (rep.iterations if min(npml) >= 4 else 2 * rep.iterations)
    sim = OracleSimulation(p.omega, p.eps, p.dl, npml, "Ez")
    for q in range(b.shape[0]):
        assert relerr(x[q].cpu().numpy(), solver_direct(sim.A, b[q]).reshape(N)) < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c1_simulation_matches_golden_and_oracle(precision):
    p = C.c1_point()
    g = np.load(os.path.join(GOLDEN, 'c1_field.npz'))
    sim = Simulation(p.omega, p.eps, p.dl, p.npml, 'Ez')
    sim.src[60, 60] = 1
    Hx, Hy, Ez = sim.solve_fields(precision=precision)
    assert Ez.dtype == np.complex128 and Ez.shape == (120, 120)
    assert relerr(Ez, g['Ez']) <= 1e-10
    ref = OracleSimulation(p.omega, p.eps, p.dl, p.npml, 'Ez')
    ref.src[60, 60] = 1
    rHx, rHy, rEz = ref.solve_fields()
    assert relerr(Hx, rHx) <= 1e-10 and relerr(Hy, rHy) <= 1e-10
    assert sim.fields['Ez'] is Ez


def test_hz_simulation_fields_match_oracle():
    rng = np.random.default_rng(3)
    N = (48, 56)
    eps = 1 + 11.25 * (rng.random(N) < 0.5)
    sim = Simulation(OMEGA, eps, 0.02, [8, 8], 'Hz')
    ref = OracleSimulation(OMEGA, eps, 0.02, [8, 8], 'Hz')
    for s in (sim, ref):
        s.src[20, 30] = 1
        s.src[10, 5] = -2
    Ex, Ey, Hz = sim.solve_fields()
    rEx, rEy, rHz = ref.solve_fields()
    for a, b in ((Ex, rEx), (Ey, rEy), (Hz, rHz)):
        assert relerr(a, b) <= 1e-10


def test_zero_source_shortcut_and_resident_factor():
    sim = Simulation(OMEGA, np.ones((40, 40)), 0.02, [8, 8], 'Ez')
    Hx, Hy, Ez = sim.solve_fields()
    assert not Ez.any() and Ez.dtype == np.complex128
    assert sim.operator.factors == {}          # no factor for a zero RHS
    sim.src[20, 20] = 1
    sim.solve_fields()
    fac = sim.operator.factors['fp64'][0]
    sim.src[10, 10] = 1
    sim.solve_fields()
    assert sim.operator.factors['fp64'][0] is fac   # reused
    sim.eps_r = np.full((40, 40), 2.0)
    assert sim.operator.factors == {}               # invalidated by the setter
    with pytest.raises(ValueError):
        sim.solve_fields(solver='umfpack')


@pytest.mark.parametrize("pol", ["Ez", "Hz"])
def test_transpose_solve_matches_csr_transpose(pol):
    rng = np.random.default_rng(4)
    N = (44, 38)
    eps = 1 + 11.25 * (rng.random(N) < 0.5)
    sim = OracleSimulation(OMEGA, eps, 0.02, [8, 7], pol)
    op = DeviceOperator(N, OMEGA, 0.02, [8, 7], pol, 1e-6, DEV)
    op.set_eps(eps)
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    x, rep = op.solve(torch.tensor(b, device=DEV), tol=1e-13, transpose=True)
    xref = solver_direct(sim.A.T.tocsr(), b.reshape(-1)).reshape(N)
    assert relerr(x.cpu().numpy(), xref) < 1e-10


def test_born_matches_golden_trajectory():
    g = np.load(os.path.join(GOLDEN, 'born_small.npz'))
    N = (64, 64)
    eps = np.ones(N)
    eps[24:40, 24:40] = 12.25
    region = np.zeros(N)
    region[24:40, 24:40] = 1
    sim = Simulation(OMEGA, eps, 0.02, [10, 10], 'Ez')
    sim.src[12, 24:40] = 40.0
    sim.add_nl(2.8e-18, region)
    Hx, Hy, Ez, conv = sim.solve_fields_nl(solver_nl='born', conv_threshold=1e-10)
    assert relerr(Ez, g['Ez']) <= 1e-10
    nsteps = int(np.count_nonzero(g['conv']))
    assert np.count_nonzero(conv) == nsteps
    # the per-step changes follow the reference trajectory until round-off
    c = conv.ravel()[:nsteps - 1]
    np.testing.assert_allclose(c, g['conv'][:nsteps - 1], rtol=1e-4)
    np.testing.assert_allclose(sim.eps_nl, g['eps_nl'], rtol=1e-9, atol=1e-14)
    # the linear operator is restored afterwards
    _, _, Elin = sim.solve_fields()
    ref = OracleSimulation(OMEGA, eps, 0.02, [10, 10], 'Ez')
    ref.src[12, 24:40] = 40.0
    _, _, rElin = ref.solve_fields()
    assert relerr(Elin, rElin) <= 1e-10


@pytest.mark.parametrize("amp", [40.0, 80.0])
def test_newton_matches_oracle_newton(amp):
    """Newton's method on the device (J11 = A with the Kerr term doubled,
    the conjugate coupling J12 refined around it) against the oracle's
    newton_solve (its real 2N system by SuperLU): same step count, the same
    per-step changes until round-off, the same fields, and the default
    solver_nl of solve_fields_nl, as in the reference."""
    N = (64, 64)
    eps = np.ones(N)
    eps[24:40, 24:40] = 12.25
    region = np.zeros(N)
    region[24:40, 24:40] = 1
    sim = Simulation(OMEGA, eps, 0.02, [10, 10], 'Ez')
    sim.src[12, 24:40] = amp
    sim.add_nl(2.8e-18, region)
    Hx, Hy, Ez, conv = sim.solve_fields_nl(conv_threshold=1e-10)
    ref = OracleSimulation(OMEGA, eps, 0.02, [10, 10], 'Ez')
    ref.src[12, 24:40] = amp
    ref.add_nl(2.8e-18, region)
    rHx, rHy, rEz, rconv = ref.newton_solve(conv_threshold=1e-10)
    n = int(np.count_nonzero(rconv))
    assert np.count_nonzero(conv) == n and n <= 5
    np.testing.assert_allclose(conv.ravel()[:n - 1], rconv.ravel()[:n - 1], rtol=1e-4)
    assert relerr(Ez, rEz) <= 1e-10
    assert relerr(Hx, rHx) <= 1e-10 and relerr(Hy, rHy) <= 1e-10
    np.testing.assert_allclose(sim.eps_nl, ref.eps_nl, rtol=1e-9, atol=1e-14)
    # the same fixed point as Born
    _, _, Eb, _ = sim.solve_fields_nl(solver_nl='born', conv_threshold=1e-12)
    assert relerr(Ez, Eb) <= 1e-9


def test_adjoint_gradient_matches_oracle():
    from fdfdpy_amd.adjoint import FieldSolver
    rng = np.random.default_rng(2)
    N = (40, 36)
    eps = np.ones(N)
    eps[14:26, 12:24] = rng.uniform(1, 12.25, size=(12, 12))
    probe = np.zeros(N)
    probe[32, 14:22] = 1
    src = np.zeros(N)
    src[6, 12:24] = 1
    solver = FieldSolver(N, OMEGA, 0.02, [8, 8], device=DEV)
    e = torch.tensor(eps, device=DEV, requires_grad=True)
    Ez = solver(e, torch.tensor(src, device=DEV))
    J = (Ez.abs() ** 2 * torch.tensor(probe, device=DEV)).sum()
    J.backward()
    ref = OracleSimulation(OMEGA, eps, 0.02, [8, 8], 'Ez')
    ref.src = src.copy()
    Jr, gr = adjoint_gradient(ref, probe)
    assert abs(J.item() - Jr) <= 1e-10 * Jr
    g = e.grad.cpu().numpy()
    assert relerr(g, gr) <= 1e-8


@pytest.mark.parametrize("precision", ["fp64", "c64"])
def test_pivoted_tile_factor_matches_oracle(precision):
    """The fallback factor (partial pivoting inside the pivot tiles)."""
    rng = np.random.default_rng(8)
    N = (50, 70)
    eps = 1 + 11.25 * (rng.random(N) < 0.5)
    sim = OracleSimulation(OMEGA, eps, 0.02, [9, 9], 'Hz')
    op = DeviceOperator(N, OMEGA, 0.02, [9, 9], 'Hz', 1e-6, DEV)
    op.set_eps(eps)
    op.factor(precision, pivot=True)
    assert op.factors[precision][2] is True
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    x, rep = op.solve(torch.tensor(b, device=DEV), precision=precision, tol=1e-13)
    assert rep.converged and not rep.fell_back
    xref = solver_direct(sim.A, b.reshape(-1)).reshape(N)
    assert relerr(x.cpu().numpy(), xref) < 1e-10


@pytest.mark.parametrize("precision,n", [("c64", 200), ("fp64", 96), ("c64", 128)])
def test_persistent_and_stepwise_factors_identical(precision, n):
    """The single-launch factorization equals the launch-per-step one bit for
    bit (one chain over all blocks, the persistent kernel's order)."""
    p = C.c3_request(77, n=n)
    ops = []
    lib = L.lib()
    prev = lib.fdfd_set_step_symmetry(0)   # full steps: the persistent kernel's arithmetic
    prev_tw = lib.fdfd_set_twist(0)
    try:
        for stepwise in (False, True):
            op = DeviceOperator(p.shape, p.omega, p.dl, p.npml, p.pol, p.L0, DEV)
            op.set_eps(p.eps)
            op.factor(precision, stepwise=stepwise)
            ops.append(op)
        torch.cuda.synchronize()
    finally:
        lib.fdfd_set_step_symmetry(prev)
        lib.fdfd_set_twist(prev_tw)
    assert torch.equal(ops[0].factors[precision][0], ops[1].factors[precision][0])


@pytest.mark.parametrize("precision", ["c64", "fp64"])
def test_twisted_factor_against_one_chain_and_oracle(precision):
    """The twisted factorization (top chain 0..c-1, bottom chain m-1..c+1,
    both in the same step launches, then block c) keeps the top blocks of the
    one-chain factorization bit for bit; its last block is T_{m-1}^-1 (the
    bottom chain starts there); the twisted sweeps then solve to the oracle's
    field, as the one-chain factor and sweeps do."""
    lib = L.lib()
    p = C.c3_request(21, n=160)
    m, n = p.shape
    npad = ((n + 31) // 32) * 32
    dt = torch.float32 if precision == "c64" else torch.float64
    rng = np.random.default_rng(21)
    b = rng.standard_normal((2,) + p.shape) + 1j * rng.standard_normal((2,) + p.shape)
    sim = OracleSimulation(p.omega, p.eps, p.dl, p.npml, p.pol, p.L0)
    blocks, xs = {}, {}
    for twist in (1, 0):
        prev = lib.fdfd_set_twist(twist)
        try:
            op = DeviceOperator(p.shape, p.omega, p.dl, p.npml, p.pol, p.L0, DEV)
            assert not op.transposed   # the oracle's block order
            op.set_eps(p.eps)
            c = lib.fdfd_twist_block(ctypes.byref(op.grid))
            assert c == ((m + 1) // 2 if twist else m - 1), c
            op.factor(precision, stepwise=True)
            blocks[twist] = (c, op.factors[precision][0].view(dt)[: m * 2 * npad * npad].clone())
            x, rep = op.solve(torch.tensor(b, device=DEV), precision=precision, tol=1e-12)
            print(f"twist {twist}: c = {c}, {rep.sweeps} sweeps, residual {rep.max_rel_residual:.2e}")
            assert rep.converged and not rep.fell_back
            xs[twist] = x.cpu().numpy()
        finally:
            lib.fdfd_set_twist(prev)
    c = blocks[1][0]
    top = c * 2 * npad * npad
    assert torch.equal(blocks[1][1][:top], blocks[0][1][:top])
    # the bottom chain's first block: G_{m-1} = T_{m-1}^-1
    A = sim.A.tocsr()
    Tl = A[(m - 1) * n:, (m - 1) * n:].toarray()
    g = blocks[1][1][(m - 1) * 2 * npad * npad:].double().cpu().numpy().reshape(2, npad, npad)
    Gl = g[0, :n, :n] + 1j * g[1, :n, :n]
    d = np.linalg.norm(Gl - np.linalg.inv(Tl)) / np.linalg.norm(Gl)
    assert d < (1e-4 if precision == "c64" else 1e-10), d
    for q in range(2):
        xref = solver_direct(A, b[q].reshape(-1)).reshape(p.shape)
        for twist in (1, 0):
            assert relerr(xs[twist][q], xref) < 1e-10


@pytest.mark.parametrize("precision,n,pol,npml", [("c64", 200, "Ez", None), ("fp64", 96, "Hz", None),
                                                   ("c64", 128, "Hz", [12, 0]), ("fp64", 160, "Ez", [0, 14])])
def test_symmetric_steps_match_full_steps(precision, n, pol, npml):
    """Symmetric steps (upper tiles only, the rest through D A = (D A)^T)
    give the full steps' block inverses to the factor's precision, and the
    solve through them meets the oracle."""
    lib = L.lib()
    p = C.c3_request(78, n=n)
    npml = p.npml if npml is None else npml
    ops = []
    for sym in (0, 1):
        prev = lib.fdfd_set_step_symmetry(sym)
        try:
            op = DeviceOperator(p.shape, p.omega, p.dl, npml, pol, p.L0, DEV)
            op.set_eps(p.eps)
            op.factor(precision, stepwise=True)
            ops.append(op)
            torch.cuda.synchronize()
        finally:
            lib.fdfd_set_step_symmetry(prev)
    # the block inverses G_k lead the factor buffer: [m][2][np][np] (square grid)
    dt = torch.float32 if precision == "c64" else torch.float64
    m, npad = p.shape[0], ((p.shape[1] + 31) // 32) * 32
    full, sym = (o.factors[precision][0].view(dt)[: m * 2 * npad * npad].double() for o in ops)
    d = ((sym - full).abs().max() / full.abs().max()).item()
    assert d < (1e-3 if precision == "c64" else 1e-9), d
    rng = np.random.default_rng(5)
    b = rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape)
    x, rep = ops[1].solve(torch.tensor(b, device=DEV), precision=precision, tol=1e-12)
    assert rep.converged and not rep.fell_back
    sim = OracleSimulation(p.omega, p.eps, p.dl, npml, pol, p.L0)
    xref = solver_direct(sim.A, b.reshape(-1)).reshape(p.shape)
    assert relerr(x.cpu().numpy(), xref) < 1e-10


@pytest.mark.parametrize("N,twist", [((64, 96), 0), ((128, 64), 1)])
def test_bounded_waits_fall_back_when_forced(N, twist):
    """With a one-poll spin limit every cross-workgroup wait gives up at once:
    the persistent factor and sweeps must fall back and still be right (one
    chain; twisted: the per-block sweeps of both chains and the twist block)."""
    lib = L.lib()
    rng = np.random.default_rng(12)
    eps = 1 + 11.25 * (rng.random(N) < 0.5)
    sim = OracleSimulation(OMEGA, eps, 0.02, [10, 10], 'Ez')
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    xref = solver_direct(sim.A, b.reshape(-1)).reshape(N)
    lib.fdfd_set_spin_cap(1)
    prev = lib.fdfd_set_twist(twist)
    try:
        op = DeviceOperator(N, OMEGA, 0.02, [10, 10], 'Ez', 1e-6, DEV)
        c = lib.fdfd_twist_block(ctypes.byref(op.grid))
        assert (c < op.grid.nx - 1) == bool(twist), c
        op.set_eps(eps)
        x, rep = op.solve(torch.tensor(b, device=DEV), precision="c64", tol=1e-12)
    finally:
        lib.fdfd_set_spin_cap(0)
        lib.fdfd_set_twist(prev)
    assert rep.converged
    assert relerr(x.cpu().numpy(), xref) < 1e-10


@pytest.mark.parametrize("shape,npml,pol", [((50, 70), [9, 9], "Ez"), ((40, 90), [8, 8], "Hz"),
                                            ((64, 48), [0, 8], "Ez")])
def test_host_csr_export_matches_oracle(shape, npml, pol):
    """Simulation.A (copied from the device bands, in the caller's
    orientation also when the solver runs transposed) and Simulation.derivs
    against the oracle's construct_A, and A again after eps_r changes."""
    rng = np.random.default_rng(21)
    eps = 1 + 11.25 * (rng.random(shape) < 0.5)
    sim = Simulation(OMEGA, eps, 0.02, npml, pol)
    ref = OracleSimulation(OMEGA, eps, 0.02, npml, pol)
    assert abs(sim.A - ref.A).max() <= 1e-13 * abs(ref.A).max()
    for k in ('Dxf', 'Dxb', 'Dyf', 'Dyb'):
        assert abs(sim.derivs[k] - ref.derivs[k]).max() <= 1e-14 * abs(ref.derivs[k]).max()
    sim.eps_r = eps * 1.5
    ref2 = OracleSimulation(OMEGA, eps * 1.5, 0.02, npml, pol)
    assert abs(sim.A - ref2.A).max() <= 1e-13 * abs(ref2.A).max()


@pytest.mark.parametrize("pol", ["Ez", "Hz"])
def test_flux_probe_on_device_matches_host_and_oracle(pol):
    """flux_probe through fdfd_flux (device-resident fields) against the host
    sum and the oracle's flux_probe (simulation.py:260-315); a line whose
    window leaves the grid takes the host sum, as numpy slices it."""
    rng = np.random.default_rng(31)
    N = (60, 50)
    eps = 1 + 11.25 * (rng.random(N) < 0.3)
    src = np.zeros(N)
    src[30, 20] = 1
    sim_t = Simulation(OMEGA, eps, 0.02, [10, 10], pol)
    sim_t.src = torch.tensor(src, device=DEV)
    sim_t.solve_fields()
    ref = OracleSimulation(OMEGA, eps, 0.02, [10, 10], pol)
    ref.src = src
    ref.solve_fields()
    for direction, center, width in [('x', [35, 25], 20), ('y', [28, 24], 16), ('x', [12, 30], 13), ('y', [40, 35], 9)]:
        F = sim_t.fields
        ix, iy = (([center[0], center[0] + 1], [int(center[1] - width / 2), int(center[1] + width / 2)])
                  if direction == 'x' else
                  ([int(center[0] - width / 2), int(center[0] + width / 2)], [center[1], center[1] + 1]))
        dev = sim_t._flux_on_device(F, direction, ix, iy)
        assert dev is not None
        want = ref.flux_probe(direction, center, width)
        assert abs(dev - want) <= 1e-9 * abs(want), (direction, center, dev, want)
        assert abs(sim_t.flux_probe(direction, center, width) - want) <= 1e-9 * abs(want)
    # window clipped by the grid edge: host sum with numpy's slicing
    assert sim_t._flux_on_device(sim_t.fields, 'x', [59, 60], [20, 30]) is None
    assert sim_t.flux_probe('x', [59, 25], 10) == ref.flux_probe('x', [59, 25], 10)


def test_include_nl_solves_with_the_kerr_diagonal_of_eps_nl():
    """solve_fields(include_nl=True) solves A + Anl (simulation.py:121-124)
    with Anl from sim.eps_nl, whatever the device operator held before: after
    a nonlinear solve and a linear solve in between. compute_nl leaves
    eps_nl, dnl_de, dnl_deps and the CSR Anl as the reference's does
    (simulation.py:57-68)."""
    N = (64, 64)
    eps = np.ones(N)
    eps[24:40, 24:40] = 12.25
    region = np.zeros(N)
    region[24:40, 24:40] = 1
    sim = Simulation(OMEGA, eps, 0.02, [10, 10], 'Ez')
    sim.src[12, 24:40] = 60.0
    sim.add_nl(2.8e-18, region, eps_scale=True, eps_max=12.25)
    ref = OracleSimulation(OMEGA, eps, 0.02, [10, 10], 'Ez')
    ref.src[12, 24:40] = 60.0
    ref.add_nl(2.8e-18, region, eps_scale=True, eps_max=12.25)
    _, _, Enl, _ = sim.solve_fields_nl(solver_nl='born', conv_threshold=1e-10)
    _, _, Elin = sim.solve_fields()               # the linear operator in between
    Hx, Hy, Ez = sim.solve_fields(include_nl=True)
    # the oracle's A + Anl with the same eps_nl
    ref.eps_nl = np.asarray(sim.eps_nl)
    Mn = N[0] * N[1]
    ref.Anl = sp.diags(OMEGA ** 2 * EPSILON_0 * 1e-6 * ref.eps_nl.reshape(-1), 0, shape=(Mn, Mn), format='csr')
    rHx, rHy, rEz = ref.solve_fields(include_nl=True)
    assert relerr(Ez, rEz) <= 1e-10 and relerr(Hx, rHx) <= 1e-10
    assert relerr(Ez, Elin) > 1e-6                 # the Kerr term matters here
    # compute_nl's host state for a given field against the oracle's
    e = np.asarray(Enl)
    sim.compute_nl(e)
    ref.compute_nl(e)
    np.testing.assert_allclose(sim.eps_nl, ref.eps_nl, rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(sim.dnl_de, ref.dnl_de, rtol=1e-12, atol=1e-14)
    expect_deps = 3 * 2.8e-18 / 1e-12 * region * np.abs(e) ** 2 / (12.25 - 1)
    np.testing.assert_allclose(sim.dnl_deps, expect_deps, rtol=1e-12, atol=1e-14)
    assert abs(sim.Anl - ref.Anl).max() <= 1e-9 * abs(ref.Anl).max()
