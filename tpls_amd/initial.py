"""Seeded initial conditions for two-layer channel flow.

Produces the five arrays main_ns_hybrid.F90 loads at start-up
(main_ns_hybrid.F90:380-410), in the shapes of the ``initial_*.dat`` files
(io.F90:47-73): u, v (0:maxl,0:maxm,0:maxn-2), w (0:maxl,0:maxm,0:maxn-1),
pressure and phi (0:maxl,0:maxm,0:maxn).  Halo rows j=0 and j=maxm hold the
periodic wrap.

* Base flow: the two-layer laminar (piecewise-parabolic) channel profile u(z)
  with no slip at z=0 and z=1, continuous velocity and shear stress at the
  interface z=height, viscosity mu_minus below and mu_plus above, scaled to
  max u = 1.  v = w = 0.
* Perturbation: uniform noise U(-A, A) on u, v, w in every cell with i >= 1
  (the inlet column i=0 stays clean, so u_inlet = u(0,0,:) is the exact
  profile); w stays 0 on the walls k=0 and k=maxn-1.
* Level set: the first-order signed distance (z - eta)/sqrt(1+|grad eta|^2)
  to eta = height + 0.05 * sum of 4 seeded sinusoids
  sin(2 pi (k x/Lx + l y/Ly) + theta).
  With the diffuse interface method (idm "dim") phi is the phase fraction
  0.5 + 0.5 tanh(d / (2 sqrt(2) epn)) of that distance d, the profile of
  twophase_initialisation_wave.F90:192-195.
* Pressure: zero (the solver's first initial guess is zero as well,
  main_ns_hybrid.F90:423).

Randomness comes from numpy's PCG64 seeded with ``seed``.
"""
from __future__ import annotations

import numpy as np

from .config import CaseParams


def two_layer_profile(z: np.ndarray, height: float, mu_lo: float, mu_hi: float) -> np.ndarray:
    """Solution of (mu u')' = -1, u(0) = u(1) = 0, mu piecewise constant, max 1."""
    h = height
    tau = (h * h / (2 * mu_lo) + (1 - h * h) / (2 * mu_hi)) / (h / mu_lo + (1 - h) / mu_hi)

    def prof(zz):
        lo = -zz * zz / (2 * mu_lo) + tau / mu_lo * zz
        hi = -zz * zz / (2 * mu_hi) + tau / mu_hi * zz + (1 / (2 * mu_hi) - tau / mu_hi)
        return np.where(zz <= h, lo, hi)

    zf = np.linspace(0.0, 1.0, 20001)
    return prof(z) / prof(zf).max()


def _wrap_y(a: np.ndarray) -> None:
    a[:, 0, :] = a[:, -2, :]
    a[:, -1, :] = a[:, 1, :]


def make_initial(p: CaseParams):
    """Return (u, v, w, pres, phi) as Fortran-ordered float64 arrays."""
    p.check()
    rng = np.random.default_rng(p.seed)
    maxl, maxm, maxn, dx = p.maxl, p.maxm, p.maxn, p.dx
    A = p.noise
    # u-grid K at z=(K+1/2)dz, w-grid K at z=K dz (index shift of pressure.f90:42)
    zu = (np.arange(maxn - 1) + 0.5) * dx
    prof = two_layer_profile(zu, p.height, p.mu_minus, p.mu_plus)

    u = np.zeros((maxl + 1, maxm + 1, maxn - 1), order="F")
    v = np.zeros((maxl + 1, maxm + 1, maxn - 1), order="F")
    w = np.zeros((maxl + 1, maxm + 1, maxn), order="F")
    u[:maxl, :, :] = prof[None, None, :]
    u[1:maxl, 1:maxm, :] += rng.uniform(-A, A, (maxl - 1, maxm - 1, maxn - 1))
    v[1:maxl, 1:maxm, :] += rng.uniform(-A, A, (maxl - 1, maxm - 1, maxn - 1))
    w[1:maxl, 1:maxm, 1:maxn - 1] += rng.uniform(-A, A, (maxl - 1, maxm - 1, maxn - 2))
    for a in (u, v, w):
        _wrap_y(a)

    Lx, Ly = (maxl - 1) * dx, (maxm - 1) * dx
    kx = rng.integers(1, 4, 4)
    ly = rng.integers(-2, 3, 4)
    th = rng.uniform(0.0, 2 * np.pi, 4)
    x = (np.arange(maxl + 1) - 0.5) * dx
    y = (np.arange(maxm + 1) - 0.5) * dx
    z = (np.arange(maxn + 1) - 0.5) * dx
    X, Y = np.meshgrid(x, y, indexing="ij")
    eta = np.full_like(X, p.height)
    gx = np.zeros_like(X)
    gy = np.zeros_like(X)
    for m in range(4):
        arg = 2 * np.pi * (kx[m] * X / Lx + ly[m] * Y / Ly) + th[m]
        eta += 0.05 * np.sin(arg)
        gx += 0.05 * np.cos(arg) * 2 * np.pi * kx[m] / Lx
        gy += 0.05 * np.cos(arg) * 2 * np.pi * ly[m] / Ly
    norm = np.sqrt(1.0 + gx * gx + gy * gy)
    phi = np.asfortranarray((z[None, None, :] - eta[:, :, None]) / norm[:, :, None])
    if p.idm == "dim":
        phi = np.asfortranarray(0.5 + 0.5 * np.tanh(phi / (2.0 * 2.0 ** 0.5 * p.dim_epn)))
    _wrap_y(phi)
    pres = np.zeros((maxl + 1, maxm + 1, maxn + 1), order="F")
    return u, v, w, pres, phi
