"""Host-side FP64 construction of the spectral operators of the quasi-cylindrical
PSATD solver.  Everything here runs once per Solver; the results are uploaded
to the device and stay there.

Restates (non-KxShift branch):
  * get_spectral_operators   solvers.py:717-759
  * grid set-up              solvers.py:60-110
  * DepFact, PoissFact       solvers.py:106, :169-176
  * PSATD coefficient tables solvers.py:227-279
  * static-solution tables   solvers.py:350-359
  * absorbing profile        solvers.py:612-617
"""
from __future__ import annotations

import functools
import hashlib
import threading
from collections import OrderedDict

import numpy as np
from scipy.special import jn, jn_zeros


def _key(v):
    if isinstance(v, np.ndarray):
        return ("a", v.shape, v.dtype.str, hashlib.sha1(np.ascontiguousarray(v).view(np.uint8)).hexdigest())
    return v


def _memo(fn):
    """Keep the last few results of a table builder, keyed by its arguments'
    values: simulations of one grid (the requests of a scan) share the tables
    instead of rebuilding them.  The cached arrays are read-only."""
    cache, lock = OrderedDict(), threading.Lock()

    @functools.wraps(fn)
    def wrapper(*args, **kw):
        key = tuple(_key(a) for a in args) + tuple((k, _key(v)) for k, v in sorted(kw.items()))
        with lock:
            if key in cache:
                cache.move_to_end(key)
                return cache[key]
        out = fn(*args, **kw)
        for a in (out.values() if isinstance(out, dict) else out):
            if isinstance(a, np.ndarray):
                a.setflags(write=False)
        with lock:
            cache[key] = out
            while len(cache) > 8:
                cache.popitem(last=False)
        return out
    return wrapper


def x_grid(leftX, rightX, dx, nthrds=None):
    """Nx and the x nodes (solvers.py:60-73): Nx even (and a multiple of the
    chunk count when Xchunked), nodes ending exactly at rightX."""
    if nthrds is not None:
        nxchunk = 2 * int(np.round(0.5 / nthrds / dx * (rightX - leftX)))
        Nx = nxchunk * nthrds
    else:
        Nx = int(2 * np.round(0.5 / dx * (rightX - leftX)))
    Xgrid = rightX - dx * np.arange(Nx)[::-1]
    return Nx, Xgrid


def r_grid(lengthR, dr):
    """Nkr and the radial nodes with the dr/2 offset (solvers.py:86-91)."""
    Nkr = int(np.round(lengthR / dr))
    RgridFull = dr * (np.arange(Nkr + 1) - 0.5)
    lengthR = RgridFull[-1] + dr
    return Nkr, RgridFull, lengthR


@_memo
def spectral_operators(Nx, Nkr, Nko, lengthR, kx, RgridFull, ext=0):
    """Bessel roots, DHT matrices and the m+-1 coupling matrices
    (solvers.py:717-759), in the reference's axis order before the swap of
    solvers.py:152-153: Out[r,k,m], In[k,r,m], DpS2S/DmS2S[k,k',m]."""
    Mmax = Nko + ext
    Mtot = Nko + 1 + ext
    Mmax_ext = Mmax + 1
    kr = np.zeros((Nkr, Mmax_ext + 1))
    kr_g = np.zeros((Nx, Nkr, Mtot))
    w = np.zeros((Nx, Nkr, Mtot))
    Out = np.zeros((Nkr, Nkr, Mtot))
    In = np.zeros((Nkr, Nkr, Mtot))
    DpS2S = np.zeros((Nkr, Nkr, Mtot))
    DmS2S = np.zeros((Nkr, Nkr, Mtot))
    for jm in range(0, Mmax_ext + 1):
        kr[:, jm] = jn_zeros(jm, Nkr) / lengthR
    R = RgridFull[1:, None]
    kx_g = None
    for jm in range(0, Mmax + 1):
        kr_g[:, :, jm], kx_g = np.meshgrid(kr[:, jm], kx)
        w[:, :, jm] = np.sqrt(kx_g ** 2 + kr_g[:, :, jm] ** 2)
        Out[:, :, jm] = jn(jm, R * kr[:, jm][None, :])
        In[:, :, jm] = np.linalg.inv(Out[:, :, jm])
        krp = kr[:, jm + 1][None, :]
        krm = kr[:, abs(jm - 1)][None, :]
        DpS2S[:, :, jm] = In[:, :, jm].dot(0.5 * krp * jn(jm, R * krp))
        DmS2S[:, :, jm] = In[:, :, jm].dot(0.5 * krm * jn(jm, R * krm))
    return dict(In=In, Out=Out, DpS2S=DpS2S, DmS2S=DmS2S, w=w, kr=kr, kr_g=kr_g, kx_g=kx_g)


@_memo
def psatd_tables(w, dt, space_charge=True):
    """PSATD_E / PSATD_G (solvers.py:244-279), (Nx,Nkr,M,5) with space
    charge, else (...,3)."""
    ncol = 5 if space_charge else 3
    E = np.zeros(w.shape + (ncol,), order="F")
    G = np.zeros_like(E)
    c, s = np.cos(dt * w), np.sin(dt * w)
    E[..., 0] = c
    E[..., 1] = s / w
    G[..., 0] = -w * s
    G[..., 1] = c
    E[..., 2] = -s / w
    G[..., 2] = 1 - c
    if space_charge:
        E[..., 3] = (dt * w * c - s) / w ** 3 / dt
        E[..., 4] = (s - dt * w) / w ** 3 / dt
        G[..., 3] = (1 - c - dt * w * s) / w ** 2 / dt
        G[..., 4] = (c - 1) / w ** 2 / dt
    return E, G


@_memo
def static_tables(kx_g, w, px0):
    """Complex tables of maxwell_solver_stat (solvers.py:350-359)."""
    beta0 = px0 / np.sqrt(1 + px0 ** 2)
    k = beta0 * kx_g[:, :, None]
    C1 = np.zeros(w.shape + (2,), dtype=complex, order="F")
    C2 = np.zeros(w.shape + (2,), dtype=complex, order="F")
    den = w ** 2 - k ** 2
    C1[..., 0] = 1.0j * k / den
    C1[..., 1] = -1.0 / den
    C2[..., 0] = w ** 2 / den
    C2[..., 1] = 1.0j * k / den
    return C1, C2


def dep_fact(Nx, dx, dr, kx_g, kr_g):
    """(2 pi)^2/Nx cos^2(kx dx/2) cos^2(kr dr/2) (solvers.py:169-176)."""
    filt_x = np.cos(0.5 * dx * kx_g)[:, :, None] ** 2
    return np.asfortranarray((2 * np.pi) ** 2 / Nx * filt_x * np.cos(0.5 * dr * kr_g) ** 2)


def damp_profile(Lf):
    """Absorbing profile (solvers.py:612-617)."""
    N = int(Lf)
    i = np.arange(N)
    return (i >= 0.75 * N) * (0.5 - 0.5 * np.cos(np.pi * (i - 0.75 * N) / (0.25 * N))) ** 2
