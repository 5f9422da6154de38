"""Host-side FP64 construction of the spectral operators of the quasi-cylindrical
PSATD solver.  Everything here runs once per Solver; the results are uploaded
to the device and stay there.

Restates:
  * get_spectral_operators   solvers.py:717-759 (both branches)
  * grid set-up              solvers.py:60-110
  * DepFact, PoissFact       solvers.py:106, :169-176
  * PSATD coefficient tables solvers.py:227-279
  * static-solution tables   solvers.py:350-359
  * absorbing profile        solvers.py:612-617
  * envelope (KxShift) tables: complex PSATD coefficients solvers.py:244-267,
    the anti-echo and band-pass filter of the deposition solvers.py:660-715
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
def spectral_operators_env(Nx, Nkr, Nko, lengthR, kx, RgridFull, ext=0):
    """The KxShift branch of get_spectral_operators (solvers.py:717-759): modes
    Mmin..Mmax = -Nko-ext..Nko+ext, the Bessel roots of |m| (scipy's jn_zeros
    of a negative order), arrays indexed by the mode with NumPy's wrap-around
    for negative m, and for Nko > 0 reordered to Mmin..Mmax along the last
    axis as the reference does (for Nko = 0 the order stays 0..Mmax, Mmin..-1:
    the envelope operators never read the coupling matrices of a single mode)."""
    Mmin, Mmax, Mtot = -Nko - ext, Nko + ext, 2 * Nko + 1 + 2 * ext
    Mmin_ext, Mmax_ext, Mtot_ext = Mmin - 1, Mmax + 1, Mtot + 2
    kr = np.zeros((Nkr, Mtot_ext))
    kr_g = np.zeros((Nx, Nkr, Mtot))
    w = np.zeros((Nx, Nkr, Mtot))
    Out = np.zeros((Nkr, Nkr, Mtot))
    In = np.zeros((Nkr, Nkr, Mtot))
    DpS2S = np.zeros((Nkr, Nkr, Mtot))
    DmS2S = np.zeros((Nkr, Nkr, Mtot))
    for jm in range(Mmin_ext, Mmax_ext + 1):
        kr[:, jm] = jn_zeros(jm, Nkr) / lengthR
    R = RgridFull[1:, None]
    kx_g = None
    for jm in range(Mmin, Mmax + 1):
        kr_g[:, :, jm], kx_g = np.meshgrid(kr[:, jm], kx)
        w[:, :, jm] = np.sqrt(kx_g ** 2 + kr_g[:, :, jm] ** 2)
        Out[:, :, jm] = jn(jm, R * kr[:, jm][None, :])
        In[:, :, jm] = np.linalg.inv(Out[:, :, jm])
        krp = kr[:, jm + 1][None, :]
        krm = kr[:, abs(jm - 1)][None, :]
        DpS2S[:, :, jm] = In[:, :, jm].dot(0.5 * krp * jn(jm, R * krp))
        DmS2S[:, :, jm] = In[:, :, jm].dot(0.5 * krm * jn(jm, R * krm))
    if Nko > 0:
        cat = lambda a: np.concatenate((a[..., Mmin:], a[..., :Mmax + 1]), axis=-1)   # noqa: E731
        Out, In, DpS2S, DmS2S, w, kr_g, kr = (cat(a) for a in (Out, In, DpS2S, DmS2S, w, kr_g, kr))
    return dict(In=In, Out=Out, DpS2S=DpS2S, DmS2S=DmS2S, w=w, kr=kr, kr_g=kr_g, kx_g=kx_g,
                idxM=np.arange(Mmin, Mmax + 1))


@_memo
def psatd_tables_env(w, kx_g, dt, beta=1.0):
    """Complex PSATD_E / PSATD_G of the envelope solver without space charge
    (solvers.py:236-267): the kx of the coefficients scaled by the frame
    velocity beta (CoPropagative)."""
    k = beta * kx_g[:, :, None]
    E = np.zeros(w.shape + (3,), dtype=complex, order="F")
    G = np.zeros_like(E)
    c, s = np.cos(dt * w), np.sin(dt * w)
    E[..., 0] = c
    E[..., 1] = s / w
    G[..., 0] = -w * s
    G[..., 1] = c
    with np.errstate(divide="ignore", invalid="ignore"):     # kx = 0 on the grid: 0/0 (NaN), as in the reference
        E[..., 2] = np.exp(-0.5j * k * dt) * 1j * k / (w ** 2 - k ** 2) * (1.0 - np.exp(1j * k * dt) * (w / (1j * k) * s + c))
        G[..., 2] = np.exp(-0.5j * k * dt) * w ** 2 / (w ** 2 - k ** 2) * (1.0 + np.exp(1j * k * dt) * (1j * k / w * s - c))
    return E, G


def antialias(kx_env, kx, kx0, Nx, features=()):
    """Deposition filter of the envelope solver (solvers.py:660-715): a band-pass
    on |kx_env| times, unless 'NoAntiEcho', a notch at each grid echo of the
    carrier inside the band (strength 'AntiEchoStrength', default 2; a list
    gives one per echo; <= 0 skips it).  Shape (Nx, 1, 1).  (The reference's
    echo count uses np.int, which NumPy 2 removed; int() truncates the same
    positive value.)"""
    cutafter = 0.85
    x = np.abs(kx_env) / np.abs(kx_env.max())
    filt_bandpass = ((x < cutafter) + (x >= cutafter) * np.cos(np.pi / 2 * (x - cutafter) / (1 - cutafter)) ** 2)
    filt_bandpass = filt_bandpass[:, None, None]
    filt_antialias = np.ones_like(filt_bandpass)
    if "NoAntiEcho" not in features:
        ae = features["AntiEchoStrength"] if "AntiEchoStrength" in features else 2
        num_echoes = int(np.abs(kx).max() / np.abs(kx_env).max()) + 1
        cell_echos = np.abs(kx_env).max() / kx0 * np.arange(num_echoes) - 1.0
        full_band = np.array([kx.min() / kx0, kx.max() / kx0]) - 1.0
        echo_ind = 0
        for cellecho in cell_echos:
            if full_band[0] < cellecho < full_band[1]:
                ae_loc = ae[echo_ind] if isinstance(ae, (list, tuple)) else ae
                echo_ind += 1
                if ae_loc <= 0:
                    continue
                xe = kx / kx0 - 1
                with np.errstate(divide="ignore", invalid="ignore"):     # the echo at -1 has zero width
                    notch = 1 - np.exp(-(xe - cellecho) ** 2 / (ae_loc * (cellecho + 1.0) / Nx) ** 2)
                filt_antialias = filt_antialias * notch[:, None, None]
    return filt_antialias * filt_bandpass


def dep_fact_env(Nx, dr, kr_g, filt_x):
    """(2 pi)^2/Nx filt_x cos^2(kr dr/2) (solvers.py:171-176, KxShift)."""
    return np.asfortranarray((2 * np.pi) ** 2 / Nx * filt_x * np.cos(0.5 * dr * kr_g) ** 2)


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
