"""Seeded generators of the five benchmark workloads (shared by tests and bench).

All workloads use dl = 0.02 (units of L0 = 1 um) and omega = 2 pi C_0 / lambda
with C_0 from constants.py:5.

1. ``c1_point``: Ez, 120 x 120, NPML [15, 15], lambda 1.55 um, eps 1 with a
   centred 30 x 30 square of 12.25, src[60, 60] = 1.
2. ``c2_design``: Hz, 512 x 512, NPML [20, 20], a 256 x 256 design region of
   Bernoulli(0.5) pixels in {1, 12.25} after a 3 x 3 box blur, 8 point
   sources outside the design region and the PML.
3. ``c3_stream``: 512 Ez requests, square grids from {128, ..., 768},
   NPML 15-25 scaled with the grid, lambda ~ U[1.3, 1.7] um, central
   half-size design region as in 2, 1-4 point sources, Poisson arrivals.
4. ``c4_adjoint``: Ez, 300 x 300, NPML [20, 20], design [100:200, 100:200]
   with eps ~ U[1, 12.25], line source src[40, 130:170] = 1, objective
   sum |Ez[260, 130:170]|^2, 32 finite-difference pixels.
5. ``c5_kerr``: Ez, 400 x 400, NPML [20, 20], eps 12.25 on [160:240]^2,
   chi3 = 2.8e-18 m^2/V^2 there, line source src[50, 170:230] = amp with 8
   log-uniform amplitudes in [3, 60]. (Amplitudes of 1e8-1e9 would give
   eps_nl ~ 1e12-1e14, where the Born map is not a contraction.)
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy import ndimage

from .constants import C_0

DL = 0.02
L0 = 1e-6


def omega_of(lam):
    return 2 * np.pi * C_0 / lam


def design_eps(rng, n_design, blur=True):
    """Bernoulli(0.5) pixels in {1, 12.25}, 3x3 box blur (mode='nearest')."""
    pix = np.where(rng.random((n_design, n_design)) < 0.5, 12.25, 1.0)
    if blur:
        pix = ndimage.uniform_filter(pix, size=3, mode='nearest')
    return pix


@dataclass
class Problem:
    name: str
    pol: str
    omega: float
    eps: np.ndarray
    npml: list
    srcs: np.ndarray                 # (nrhs, Nx, Ny) real source arrays
    dl: float = DL
    L0: float = L0
    extra: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.eps.shape

    def rhs(self):
        """b = 1j omega src per RHS (simulation.py:121)."""
        return self.srcs * (1j * self.omega)


def c1_point():
    N = (120, 120)
    eps = np.ones(N)
    eps[45:75, 45:75] = 12.25
    src = np.zeros((1,) + N)
    src[0, 60, 60] = 1
    return Problem('c1_point', 'Ez', omega_of(1.55e-6), eps, [15, 15], src)


def point_sources(rng, n, npml, lo, hi, k):
    """k distinct cells outside [lo, hi)^2 and outside the PML."""
    a, b = npml + 2, n - npml - 2
    side = max(0, b - a)
    inner = max(0, min(b, hi) - max(a, lo))
    if side * side - inner * inner < k:
        raise ValueError(f"no room for {k} sources outside the design region of a {n}-cell grid "
                         f"with NPML {npml}")
    cells = set()
    while len(cells) < k:
        i, j = rng.integers(npml + 2, n - npml - 2, size=2)
        if lo <= i < hi and lo <= j < hi:
            continue
        cells.add((int(i), int(j)))
    return sorted(cells)


def c2_design(seed=2, n=512, nrhs=8):
    rng = np.random.default_rng(seed)
    eps = np.ones((n, n))
    lo, hi = n // 4, 3 * n // 4
    eps[lo:hi, lo:hi] = design_eps(rng, hi - lo)
    src = np.zeros((nrhs, n, n))
    for q, (i, j) in enumerate(point_sources(rng, n, 20, lo, hi, nrhs)):
        src[q, i, j] = 1
    return Problem('c2_design', 'Hz', omega_of(1.55e-6), eps, [20, 20], src)


STREAM_SIZES = (128, 192, 256, 384, 512, 768)


def stream_npml(n):
    return int(round(15 + 10 * (n - 128) / 640))


def c3_request(seed, n=None):
    """One request of the serving stream (per-request seed)."""
    rng = np.random.default_rng(1000 + seed)
    if n is None:
        n = int(rng.choice(STREAM_SIZES))
    lam = rng.uniform(1.3e-6, 1.7e-6)
    npml = stream_npml(n)
    eps = np.ones((n, n))
    lo, hi = n // 4, 3 * n // 4
    eps[lo:hi, lo:hi] = design_eps(rng, hi - lo)
    k = int(rng.integers(1, 5))
    src = np.zeros((k, n, n))
    for q, (i, j) in enumerate(point_sources(rng, n, npml, lo, hi, k)):
        src[q, i, j] = 1
    return Problem(f'c3_req{seed}', 'Ez', omega_of(lam), eps, [npml, npml], src,
                   extra={'lambda': lam})


def c3_stream(n_requests=512, seed=3, rate=None):
    """Requests and Poisson arrival times (rate in requests/s; None = all at t=0)."""
    rng = np.random.default_rng(seed)
    if rate:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
        arrivals -= arrivals[0]
    else:
        arrivals = np.zeros(n_requests)
    return [c3_request(s) for s in range(n_requests)], arrivals


def c4_adjoint(seed=4, n=300):
    rng = np.random.default_rng(seed)
    eps = np.ones((n, n))
    eps[100:200, 100:200] = rng.uniform(1.0, 12.25, size=(100, 100))
    src = np.zeros((1, n, n))
    src[0, 40, 130:170] = 1
    probe = np.zeros((n, n))
    probe[260, 130:170] = 1
    design = np.zeros((n, n), dtype=bool)
    design[100:200, 100:200] = True
    flat = np.flatnonzero(design)
    fd_pixels = [tuple(int(v) for v in np.unravel_index(p, (n, n)))
                 for p in rng.choice(flat, size=32, replace=False)]
    return Problem('c4_adjoint', 'Ez', omega_of(1.55e-6), eps, [20, 20], src,
                   extra={'probe': probe, 'design': design, 'fd_pixels': fd_pixels, 'fd_step': 1e-4})


def c5_kerr(seed=5, n=400, n_amp=8, amp_range=(3.0, 60.0)):
    rng = np.random.default_rng(seed)
    eps = np.ones((n, n))
    eps[160:240, 160:240] = 12.25
    region = np.zeros((n, n))
    region[160:240, 160:240] = 1
    amps = np.exp(rng.uniform(np.log(amp_range[0]), np.log(amp_range[1]), size=n_amp))
    src = np.zeros((1, n, n))
    src[0, 50, 170:230] = 1
    return Problem('c5_kerr', 'Ez', omega_of(1.55e-6), eps, [20, 20], src,
                   extra={'chi3': 2.8e-18, 'nl_region': region, 'amps': amps})
