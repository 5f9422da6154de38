"""Seeded inputs shared by the CPU and GPU parity tests.

Each generator returns the positional arguments of one f2py entry point (in the
reference's calling convention), built from a fixed seed so every back end sees
the same bytes.  Sizes are small enough for the C oracle to finish in well under
a second.
"""
from __future__ import annotations

import numpy as np


class Grid:
    """A quasi-cylindrical grid laid out as the reference's Solver does
    (solvers.py:60-91): x nodes leftX + dx*i, radial nodes dr*(i-1/2)."""

    def __init__(self, nx=64, nkr=16, nko=2, dx=0.1, dr=0.3, leftX=-3.15):
        self.nx, self.nkr, self.nko, self.dx, self.dr, self.leftX = nx, nkr, nko, dx, dr, leftX
        self.nmod = nko + 1
        self.Rgrid = dr * (np.arange(nkr + 1) - 0.5)
        self.kx = 2 * np.pi * np.fft.fftfreq(nx, dx)


def rng(seed):
    return np.random.default_rng(seed)


def F(r, *shape):
    return np.asfortranarray(r.standard_normal(shape))


def C(r, *shape):
    return np.asfortranarray(r.standard_normal(shape) + 1j * r.standard_normal(shape))


def particles(g: Grid, n=4000, seed=1, margin=0.02, zero_weights=True):
    """Positions strictly inside the grid (so no particle indexes outside the
    reference's arrays), momenta of order 1, some zero weights, one particle
    exactly on the axis and one with zero momentum."""
    r = rng(seed)
    x = r.uniform(g.leftX + margin, g.leftX + (g.nx - 1) * g.dx - margin, n)
    rr = r.uniform(0.0, g.Rgrid[-1] - margin, n)
    th = r.uniform(0, 2 * np.pi, n)
    coord = np.asfortranarray(np.vstack([x, rr * np.sin(th), rr * np.cos(th)]))
    coord[1:, 0] = 0.0                       # on the axis: phase 0 (grid_deps.f90:55-59)
    mom = F(r, 3, n)
    if n > 1:
        mom[:, 1] = 0.0                      # zero momentum: skipped by dep_curr (:42)
    w = r.uniform(0.1, 1.0, n)
    if zero_weights:
        w[::37] = 0.0                        # skipped everywhere (:36, :106, :171)
    return coord, mom, np.asfortranarray(w)


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0)


def rel_max(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)
