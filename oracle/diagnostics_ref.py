"""CPU restatement of the reference's diagnostics on host arrays in its
layouts: the beam envelopes of get_beam_envelops (diagnostics.py:177-207;
coords_halfstep and momenta as (3, np), weights (np,)), the EM energy per kx
of nrg_out (diagnostics.py:109-122) with the EnergyFact of
solvers.py:142-149, the laser power of pwr_out (diagnostics.py:124-149)
through the C port's inverse transform, the density of get_dens_on_grid
(species.py:400-427), the phase-space selection of phs_out
(diagnostics.py:89-107) and the azimuthal profile intens_profO
(utils.f90:18-57).

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


def dens_on_grid(port, coords, weights, leftX, Rgrid, dx, dr, Nx, Nko):
    """species.py:400-427: get_dens_on_grid of a species on its grid (Nx x
    len(Rgrid) nodes, modes 0..Nko): the oracle port's dep_dens (ghost fold
    included) times 1/(2 pi dx dr Rgrid) per radial row, 0 where Rgrid <= 0."""
    VGrid = 2 * np.pi * dx * dr * Rgrid
    VGrid = (VGrid + (Rgrid == 0)) ** -1 * (Rgrid > 0.0)
    dens = np.zeros((Nx, Rgrid.shape[0], Nko + 1), dtype="complex", order="F")
    dens = port.dep_dens(np.asfortranarray(coords), np.asarray(weights, float), dens, leftX, Rgrid, 1.0 / dx,
                         1 / dr) * VGrid[None, :, None]
    return dens


def phase_space(coords_halfstep, momenta, weights):
    """diagnostics.py:94-103: PartsPack of the particles with
    (1 + u**2).sum(0) > 20**2."""
    indx = np.nonzero((1. + momenta ** 2).sum(0) > 20. ** 2)[0]
    return np.concatenate((coords_halfstep[:, indx], momenta[:, indx], weights[indx][None, :]), axis=0)


def intens_profo(Fld, NO):
    """utils.f90:18-57 (f2py name intens_profo): Fld(0:nx, 0:nr, -nkO:nkO, 3);
    PWR_RO(NO, 1:nr) = sum over l and ix of |sum_m phaseO(m) Fld(ix, ir, m, l)|^2
    with the phases formed by the reference's recurrences (Os(iO) =
    Os(iO-1) * e^{2 pi i/(NO-1)}; phaseO(+-k) by repeated products)."""
    nm = Fld.shape[2]
    if nm % 2 == 0:
        raise ValueError("intens_profo: the mode axis is -nkO..nkO (odd length)")
    nkO = (nm - 1) // 2
    nr = Fld.shape[1] - 1
    Os = np.zeros(NO, complex)
    Os[0] = 1.0
    phase_p = np.cos(2.0 * np.pi / (NO - 1)) + 1j * np.sin(2.0 * np.pi / (NO - 1))
    for iO in range(1, NO):
        Os[iO] = Os[iO - 1] * phase_p
    PWR = np.zeros((NO, nr))
    for iO in range(NO):
        pp, pm = Os[iO], 1.0 / Os[iO]
        phaseO = np.ones(nm, complex)
        for k in range(1, nkO + 1):
            phaseO[nkO + k] = phaseO[nkO + k - 1] * pp
            phaseO[nkO - k] = phaseO[nkO - k + 1] * pm
        s = np.einsum("m,xrml->xrl", phaseO, Fld[:, 1:, :, :])
        PWR[iO] = (np.abs(s) ** 2).sum(axis=(0, 2))
    return PWR
