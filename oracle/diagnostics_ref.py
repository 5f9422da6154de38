"""CPU restatement of three of the reference's diagnostics on host arrays in
its layouts: the beam envelopes of get_beam_envelops (diagnostics.py:177-207;
coords_halfstep and momenta as (3, np), weights (np,)), the EM energy per kx
of nrg_out (diagnostics.py:109-122) with the EnergyFact of
solvers.py:142-149, and the laser power of pwr_out (diagnostics.py:124-149)
through the C port's inverse transform.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker; the product never
imports it.
"""
from __future__ import annotations

import numpy as np


def beam_envelope(coords_halfstep, momenta, weights):
    """diagnostics.py:185-203: [[<x>, <y>, <z>], [rms x, rms y, rms z],
    [emittance x, y, z]] with the reference's uncentred emittance
    sqrt(<x^2><px^2> - <x px>^2) and its rms sqrt(<x^2> - <x>^2)."""
    x, y, z = coords_halfstep[0], coords_halfstep[1], coords_halfstep[2]
    px, py, pz = momenta[0], momenta[1], momenta[2]
    w = weights
    xyz_0 = [(x * w).sum() / w.sum(), (y * w).sum() / w.sum(), (z * w).sum() / w.sum()]
    xyz_rms = [np.sqrt((x ** 2 * w).sum() / w.sum() - xyz_0[0] ** 2),
               np.sqrt((y ** 2 * w).sum() / w.sum() - xyz_0[1] ** 2),
               np.sqrt((z ** 2 * w).sum() / w.sum() - xyz_0[2] ** 2)]
    xyz_emmit = [np.sqrt((x ** 2 * w).sum() / w.sum() * (px ** 2 * w).sum() / w.sum() - (x * px * w).sum() ** 2 / w.sum() ** 2),
                 np.sqrt((y ** 2 * w).sum() / w.sum() * (py ** 2 * w).sum() / w.sum() - (y * py * w).sum() ** 2 / w.sum() ** 2),
                 np.sqrt((z ** 2 * w).sum() / w.sum() * (pz ** 2 * w).sum() / w.sum() - (z * pz * w).sum() ** 2 / w.sum() ** 2)]
    return np.array([xyz_0, xyz_rms, xyz_emmit])


def moment_sums(coords_halfstep, momenta, weights):
    """The 13 weighted sums the envelope is formed from, in the order of
    chimera_beam_moments (products as diagnostics.py:187-203 writes them)."""
    x, y, z = coords_halfstep[0], coords_halfstep[1], coords_halfstep[2]
    px, py, pz = momenta[0], momenta[1], momenta[2]
    w = weights
    return np.array([w.sum(), (x * w).sum(), (y * w).sum(), (z * w).sum(), (x ** 2 * w).sum(), (y ** 2 * w).sum(),
                     (z ** 2 * w).sum(), (px ** 2 * w).sum(), (py ** 2 * w).sum(), (pz ** 2 * w).sum(),
                     (x * px * w).sum(), (y * py * w).sum(), (z * pz * w).sum()])


def energy_fact(idxM, kr_g, lengthR, dkx, kx_shift=False):
    """The energy integration factor of solvers.py:142-149."""
    from scipy.constants import c, e, epsilon_0, m_e
    from scipy.special import jn
    tmp_idx = (np.abs(idxM) + 1)[None, None, :]
    EF = 0.5e-2 * (m_e * c ** 2 / e) ** 2 * (4 * np.pi * epsilon_0) * lengthR ** 2 / dkx \
        * jn(tmp_idx, kr_g * lengthR) ** 2
    if not kx_shift:
        EF *= 0.5
    return EF


def nrg_out(EG_fb, EnergyFact):
    """diagnostics.py:113-116: EM energy per kx, the kx axis reordered."""
    dat = ((abs(EG_fb[:, :, :, :3]) ** 2).sum(-1) * EnergyFact).sum(-1).sum(-1)
    return np.r_[dat[dat.shape[0] // 2 + 1:], dat[:dat.shape[0] // 2 + 1]]


PwrFctr = 0.5 * 0.511e6 * 1.6022e-19 / 2.818e-13 * 2.9979e10      # diagnostics.py:24


def pwr_out(port, EG_fb, leftX, kx, OutFull, RgridFull, dr):
    """diagnostics.py:128-147 (no 'Spot'): fb_vec_out through the oracle port's
    fb_eb_out (its first three components are fb_vec_out, fb_io.f90:100-140
    and :182-228), then PwrFctr*2*dr*((|dat|^2).sum(-1).sum(-1)*Rgrid).sum(-1)."""
    nkx, nkr, M = EG_fb.shape[:3]
    nr = OutFull.shape[1]
    e = np.asfortranarray(EG_fb[..., :3])
    eb = port.fb_eb_out(np.zeros((nkx, nr + 1, M, 6), complex, order="F"), np.asfortranarray(EG_fb), e, leftX, kx,
                        OutFull)
    dat = eb[..., :3]
    return PwrFctr * 2 * dr * ((np.abs(dat) ** 2).sum(-1).sum(-1) * RgridFull[None, :]).sum(-1)
