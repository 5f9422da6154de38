"""Diagnostics computed on the device, with the interface of the reference's
Diagnostics class (diagnostics.py:25-207) for what a served simulation
returns: the beam envelopes of get_beam_envelops, formed on the host from 13
weighted sums that chimera_beam_moments reduces on the GPU, so that no
particle array leaves the device.

The reference's dump diagnostics (fields, FB fields, particles, density, EM
energy, power; diagnostics.py:33-150), which write files, are not provided.
"""
from __future__ import annotations

import numpy as np

from .device import call


class Diagnostics:
    def __init__(self, Chimera, diags=(), out_folder=None):
        self.Chimera = Chimera
        self.diags = diags
        self.out_folder = out_folder

    @staticmethod
    def beam_moment_sums(species):
        """sum w, sum x w, sum y w, sum z w, sum x^2 w, sum y^2 w, sum z^2 w,
        sum px^2 w, sum py^2 w, sum pz^2 w, sum x px w, sum y py w, sum z pz w
        over the species' live particles (x, y, z at the half step)."""
        sums = np.zeros(13)
        fn = "chimera_beam_moments_mixed" if species.mixed else "chimera_beam_moments"
        call(fn, species.ptr("coords_halfstep"), species.ptr("momenta"), species.ptr("weights"), species.ld,
             species.np, sums.ctypes.data)
        return sums

    def get_beam_envelops(self):
        """diagnostics.py:177-207: per species that is not Still, the array
        [[<x>, <y>, <z>], [rms x, rms y, rms z], [emittance x, y, z]] with the
        reference's formulas (uncentred emittance sqrt(<x^2><px^2> - <x px>^2))."""
        out = []
        for sp in self.Chimera.Particles:
            if "Still" in sp.Args["Features"]:
                continue
            out.append(self.envelope_from_sums(self.beam_moment_sums(sp)))
        return out

    @staticmethod
    def envelope_from_sums(S):
        """The envelope array of diagnostics.py:185-203 from the 13 sums."""
        W = S[0]
        xyz_0 = S[1:4] / W
        xyz_rms = np.sqrt(S[4:7] / W - xyz_0 ** 2)
        xyz_emmit = np.sqrt(S[4:7] / W * S[7:10] / W - S[10:13] ** 2 / W ** 2)
        return np.array([xyz_0, xyz_rms, xyz_emmit])
