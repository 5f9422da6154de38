"""Diagnostics computed on the device, with the interface of the reference's
Diagnostics class (diagnostics.py:25-207):

* get_beam_envelops: the beam envelopes, formed on the host from 13 weighted
  sums that chimera_beam_moments reduces on the GPU, so that no particle
  array leaves the device;
* nrg_out: the EM energy per kx (chimera_field_energy), returned ('Return'
  in the diagnostic's Features) or appended to nrg_<solver>.txt as the
  reference does;
* pwr_out: the laser power per x (the inverse transform of fb_vec_out, then
  chimera_line_power), returned or appended to pwrX_<solver>.txt; the 'Spot'
  intensity profile is not provided;
* do_diags runs the 'EnergyEM' and 'Power' entries of the list.

The dump diagnostics of the reference (fields, FB fields, particles, density;
diagnostics.py:46-108) are not provided.
"""
from __future__ import annotations

import os

import numpy as np

from .device import DeviceArray, call

PwrFctr = 0.5 * 0.511e6 * 1.6022e-19 / 2.818e-13 * 2.9979e10      # diagnostics.py:24


class Diagnostics:
    def __init__(self, Chimera, diags=(), out_folder=None):
        self.Chimera = Chimera
        self.diags = diags
        self.out_folder = out_folder
        if out_folder is not None:
            # created if absent (the reference deletes and recreates it)
            os.makedirs(out_folder, exist_ok=True)

    def do_diags(self, i):
        """diagnostics.py:34-46 for the types provided here ('EnergyEM')."""
        for diag in self.diags:
            if np.mod(i, diag["Step"]) != 0:
                continue
            diag.setdefault("Features", {})
            if diag["Type"] == "EnergyEM":
                self.nrg_out(diag)
            elif diag["Type"] == "Power":
                self.pwr_out(diag)
            else:
                raise NotImplementedError(f"diagnostic type {diag['Type']!r} is not provided on the device")

    @staticmethod
    def energy_fact_table(A):
        """EnergyFact of solvers.py:142-149 from a solver's Args (host array)."""
        from scipy.constants import c, e, epsilon_0, m_e
        from scipy.special import jn
        tmp_idx = (np.abs(A["idxM"]) + 1)[None, None, :]
        EF = 0.5e-2 * (m_e * c ** 2 / e) ** 2 * (4 * np.pi * epsilon_0) * A["lengthR"] ** 2 / A["dkx"] \
            * jn(tmp_idx, A["kr_g"] * A["lengthR"]) ** 2
        if "KxShift" not in A:
            EF *= 0.5
        return np.asfortranarray(EF)

    @classmethod
    def energy_fact(cls, solver):
        """EnergyFact kept on the device with the solver."""
        if "EnergyFact" not in solver.dev:
            solver.dev["EnergyFact"] = DeviceArray.from_numpy(cls.energy_fact_table(solver.Args), np.float64)
        return solver.dev["EnergyFact"]

    def nrg_out(self, diag):
        """diagnostics.py:109-122."""
        ret = []
        for jj, sol in enumerate(self.Chimera.Solvers):
            Nx, Nkr, M = sol.n_fb
            dat = np.zeros(Nx)
            ef = self.energy_fact(sol)
            call("chimera_field_energy_c8" if sol.mixed else "chimera_field_energy", sol.Data["EG_fb"].ptr, ef.ptr,
                 Nx, Nkr, M, dat.ctypes.data)
            dat = np.r_[dat[dat.shape[0] // 2 + 1:], dat[:dat.shape[0] // 2 + 1]]
            if "Return" in diag["Features"]:
                ret.append(dat.copy())
            else:
                if self.out_folder is None:
                    raise ValueError("nrg_out writes nrg_<n>.txt into out_folder: give one, or 'Return' in Features")
                with open(os.path.join(self.out_folder, f"nrg_{jj}.txt"), mode="a") as fout:
                    np.savetxt(fout, dat[None, :])
        if "Return" in diag["Features"]:
            return ret

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

    @staticmethod
    def _table(solver, key, host, dtype):
        if key not in solver.dev:
            solver.dev[key] = DeviceArray.from_numpy(np.asfortranarray(host), dtype)
        return solver.dev[key]

    def pwr_out(self, diag):
        """diagnostics.py:124-149 without the 'Spot' profile: the field of
        fb_vec_out (fb_io.f90:100-140; the E part of the inverse transform with
        the full radial grid, ghost row zero) reduced on the device to
        PwrFctr*2*dr * sum_r Rgrid[r] sum_m sum_c |E|^2 per x."""
        if "Spot" in diag["Features"]:
            raise NotImplementedError("the 'Spot' intensity profile (intens_profo) is not provided on the device")
        ret = []
        for jj, sol in enumerate(self.Chimera.Solvers):
            A = sol.Args
            Nx, Nkr, M = sol.n_fb
            nr = A["OutFull"].shape[1]
            out_full = self._table(sol, "OutFull", A["OutFull"], sol.rt)
            rgrid = self._table(sol, "RgridFull", A["RgridFull"], np.float64)
            vec = sol.dev.get("PwrVec")
            if vec is None:
                vec = sol.dev["PwrVec"] = DeviceArray((Nx, nr + 1, M, 6), sol.ct)
            EG = sol.Data["EG_fb"]
            # E in both halves of the transform's input: the first three output components are fb_vec_out
            call(sol.fn("chimera_fb_inverse"), vec.ptr, EG.ptr, EG.ptr, float(A["leftX"]), sol.dev["kx"].ptr,
                 out_full.ptr, Nx, nr, M, Nkr, 1, None)
            dat = np.zeros(Nx)
            call("chimera_line_power_c8" if sol.mixed else "chimera_line_power", vec.ptr, rgrid.ptr, Nx, nr, M,
                 dat.ctypes.data)
            dat = PwrFctr * 2 * A["dr"] * dat
            if "Return" in diag["Features"]:
                ret.append(dat)
            else:
                if self.out_folder is None:
                    raise ValueError("pwr_out writes pwrX_<n>.txt into out_folder: give one, or 'Return' in Features")
                with open(os.path.join(self.out_folder, f"pwrX_{jj}.txt"), mode="a") as fout:
                    np.savetxt(fout, dat[None])
        if "Return" in diag["Features"]:
            return ret
