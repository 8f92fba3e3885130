"""The reference's own result checks, run on the device path.

- tests/test_flux.py:11-37: the flux of a modal source through a probe line
  does not depend on the mesh (300 x 100 at dl = 0.01 against 600 x 200 at
  dl = 0.005), rtol 1e-3.
- tests/test_nonlinear_solvers.py:12-58: Newton's method and the Born
  iteration reach the same Kerr field with eps_scale=True (the branch of
  nonlinearity.py:19-26). The reference compares ``fields['Ez']``, which
  neither solver writes (both fill ``fields_nl``), so it compares the linear
  field with itself; here the nonlinear fields are compared.
- tests/test_chi3.py:9-79: the self-phase modulation of a waveguide mode
  against the analytic nonlinear phase gamma P L. The reference test calls a
  solve_fields_nl signature that no longer exists and reads the linear
  fields; here the Kerr term is added with add_nl (chi3 / L0^2, simulation.py:
  70-73) and the phase and power are read from the nonlinear fields.
"""

import numpy as np
import pytest

from fdfdpy_amd import Simulation

pytestmark = pytest.mark.gpu


def test_flux_resolution_consistency():
    omega = 2 * np.pi * 200e12
    eps1 = np.ones((300, 100))
    eps1[:, 40:60] = 12.25
    s1 = Simulation(omega, eps1, 0.01, [15, 15], 'Ez')
    s1.add_mode(3.5, 'x', [20, 50], 60, scale=1)
    s1.setup_modes()
    s1.solve_fields()
    flux1 = s1.flux_probe('x', [150, 50], 60)
    eps2 = np.ones((600, 200))
    eps2[:, 80:120] = 12.25
    s2 = Simulation(omega, eps2, 0.005, [15, 15], 'Ez')
    s2.add_mode(3.5, 'x', [20, 100], 120, scale=1)
    s2.setup_modes()
    s2.solve_fields()
    flux2 = s2.flux_probe('x', [300, 100], 120)
    print(f"flux 300x100 {flux1:.9e}, 600x200 {flux2:.9e}, relative difference {abs(flux1 - flux2) / abs(flux2):.2e}")
    np.testing.assert_allclose(flux1, flux2, rtol=1e-3)


def _kerr_waveguide(eps_scale):
    """The waveguide of test_chi3.py / test_nonlinear_solvers.py: n0 = 3.4,
    1 um wide, 5 um long, chi3 = 2.8e-18 over the central 4 um."""
    n0, dl, chi3 = 3.4, 0.01, 2.8e-18
    width, L, L_chi3 = 1, 5, 4
    wv, lv = int(width / dl), int(L_chi3 / dl)
    Nx, Ny = int(L / dl), int(3.5 * width / dl)
    eps = np.ones((Nx, Ny))
    eps[:, int(Ny / 2 - wv / 2):int(Ny / 2 + wv / 2)] = np.square(n0)
    region = np.zeros(eps.shape)
    region[int(Nx / 2 - lv / 2):int(Nx / 2 + lv / 2), int(Ny / 2 - wv / 2):int(Ny / 2 + wv / 2)] = 1
    sim = Simulation(2 * np.pi * 200e12, eps, dl, [15, 15], 'Ez')
    sim.add_mode(n0, 'x', [17, int(Ny / 2)], wv * 3)
    sim.setup_modes()
    if eps_scale:
        sim.add_nl(chi3, region, eps_scale=True, eps_max=np.max(eps))
    else:
        sim.add_nl(chi3, region)
    return sim, dict(n0=n0, dl=dl, chi3=chi3, wv=wv, lv=lv, Nx=Nx, Ny=Ny)


def test_newton_equals_born_with_eps_scale():
    sim, g = _kerr_waveguide(eps_scale=True)
    for srcval in np.logspace(1, 3, 3):
        sim.setup_modes()
        sim.src = sim.src * srcval
        sim.fields = {k: None for k in sim.fields}
        _, _, E_newton, conv_n = sim.solve_fields_nl(solver_nl='newton')
        E_newton = np.array(sim.fields_nl['Ez'])
        _, _, E_born, conv_b = sim.solve_fields_nl(solver_nl='born')
        E_born = np.array(sim.fields_nl['Ez'])
        assert np.count_nonzero(sim.dnl_deps) > 0      # the eps_scale branch: d eps_nl / d eps_r != 0
        diff = np.linalg.norm(E_newton - E_born) / np.linalg.norm(E_born)
        print(f"srcval {srcval:g}: Newton {np.count_nonzero(conv_n)} steps, Born {np.count_nonzero(conv_b)} "
              f"steps, |E_newton - E_born| / |E_born| = {diff:.2e}")
        np.testing.assert_allclose(E_newton, E_born, rtol=1e-3, atol=1e-6 * np.abs(E_born).max())
        assert diff <= 1e-8


def test_chi3_self_phase_modulation_matches_analytic():
    sim, g = _kerr_waveguide(eps_scale=False)
    Nx, Ny, n0, dl, chi3 = g['Nx'], g['Ny'], g['n0'], g['dl'], g['chi3']
    sim.solve_fields()
    T_linear = sim.fields['Ez'][Nx - 20, int(Ny / 2)] / sim.fields['Ez'][20, int(Ny / 2)]
    T_vec, pwr_vec = [], []
    for srcval in np.logspace(1, 3, 3):
        sim.setup_modes()
        sim.src = sim.src * srcval
        sim.fields = {k: None for k in sim.fields}    # Newton starts from this source's linear field
        sim.solve_fields_nl(solver_nl='newton', Estart=None)
        E = np.array(sim.fields_nl['Ez'])
        T_vec.append(E[Nx - 20, int(Ny / 2)] / E[20, int(Ny / 2)])
        pwr_vec.append(sim.flux_probe('x', [Nx - 20, int(Ny / 2)], g['wv'] * 3, nl=True))
    T_vec, pwr_vec = np.array(T_vec), np.array(pwr_vec)
    # test_chi3.py:63-75: n2 from chi3, gamma = k0 n2 / Aeff, phase gamma P L
    n2 = 12 * np.square(np.pi) * chi3 * 1e4 / np.square(n0)
    n2 *= 1e-4 / np.square(sim.L0)
    width = dl * g['wv']
    height = width
    Aeff = width * height
    L = dl * g['lv']
    gamma_spm = (sim.omega / 299792458 * sim.L0) * n2 / Aeff
    Phi_fdfd = -np.unwrap(np.angle(T_vec) - np.angle(T_linear)) / np.pi
    Phi_analytic = (pwr_vec * height) * L * gamma_spm / np.pi
    print("Phi_fdfd", Phi_fdfd, "Phi_analytic", Phi_analytic, "ratio", Phi_fdfd / Phi_analytic)
    np.testing.assert_allclose(Phi_fdfd, Phi_analytic, rtol=1e-3)
