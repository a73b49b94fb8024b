"""Seeded Fortran-shaped test fields for the per-routine parity tests.

Arrays follow the reference shapes of main_ns_hybrid.F90:416-549 on a single
rank (sx=1, ex=maxl-1, sy=1, ey=maxm-1): p-grid (0:maxl,0:maxm,0:maxn),
u/v (0:maxl,0:maxm,0:maxn-2), w (0:maxl,0:maxm,0:maxn-1), the augmented
viscosity (-1:maxl+1,-1:maxm+1,0:maxn) and phi_augaug
(-2:maxl+2,-2:maxm+2,0:maxn).  Halo rows hold the periodic wrap of the
interior, as after the reference's exchange2d* calls.
"""
import numpy as np


def wrap_y(a, hw):
    """Fill hw halo rows on each side in y with the periodic wrap (in place)."""
    ny = a.shape[1] - 2 * hw          # interior rows sy..ey
    for d in range(1, hw + 1):
        a[:, hw - d, :] = a[:, hw - d + ny, :]
        a[:, hw + ny - 1 + d, :] = a[:, hw + d - 1, :]
    return a


def shapes(maxl, maxm, maxn):
    return {
        "p": (maxl + 1, maxm + 1, maxn + 1),
        "uv": (maxl + 1, maxm + 1, maxn - 1),
        "w": (maxl + 1, maxm + 1, maxn),
        "aug": (maxl + 3, maxm + 3, maxn + 1),
        "augaug": (maxl + 5, maxm + 5, maxn + 1),
    }


def rand(rng, shape, scale=1.0, offset=0.0, hw=1):
    a = np.asfortranarray(offset + scale * rng.standard_normal(shape))
    return wrap_y(a, hw)


def level_set(rng, maxl, maxm, maxn, dx, height=0.5, amp=0.05, noise=0.0, hw=1):
    """phi = z - eta(x, y) on the p-grid, cell centres ((i-1/2)dx, ...)."""
    i = np.arange(maxl + 1 + 2 * (hw - 1)) - (hw - 1)
    j = np.arange(maxm + 1 + 2 * (hw - 1)) - (hw - 1)
    k = np.arange(maxn + 1)
    X, Y, Z = np.meshgrid((i - 0.5) * dx, (j - 0.5) * dx, (k - 0.5) * dx, indexing="ij")
    Lx, Ly = (maxl - 1) * dx, (maxm - 1) * dx
    th = rng.uniform(0, 2 * np.pi, 3)
    eta = height + amp * (np.sin(2 * np.pi * X / Lx + th[0]) + np.sin(2 * np.pi * Y / Ly + th[1])
                          + 0.5 * np.sin(2 * np.pi * (X / Lx + Y / Ly) + th[2]))
    phi = Z - eta
    if noise:
        phi = phi + noise * rng.standard_normal(phi.shape)
    return wrap_y(np.asfortranarray(phi), hw)
