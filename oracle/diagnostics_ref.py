"""CPU restatement of the beam envelopes of the reference's diagnostics
(get_beam_envelops, diagnostics.py:177-207), on host arrays in the
reference's layout (coords_halfstep and momenta as (3, np), weights (np,)).

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
