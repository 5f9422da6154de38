"""Diagnostics with the interface of the reference's Diagnostics class
(diagnostics.py:25-207), computed on the device:

* fld_out / fldfb_out: the real-space fields EB (ghost row dropped) and the
  FB-space fields EG_fb, copied from the device;
* dns_out: each pushed species' density on the grid (Specie.get_dens_on_grid,
  deposited on the device);
* phs_out: the phase space of the particles above the LPA energy filter,
  selected and packed on the device (chimera_phs_select);
* nrg_out: the EM energy per kx (chimera_field_energy);
* pwr_out: the laser power per x (the inverse transform of fb_vec_out, then
  chimera_line_power) and, with 'Spot', the azimuthal intensity profile
  intens_profO (chimera_intens_profo);
* get_spot_cartesian: that profile interpolated onto a Cartesian grid
  (scipy griddata on the host, as the reference);
* get_beam_envelops: the beam envelopes from 13 weighted sums reduced on the
  GPU, so that no particle array leaves the device.

Each dump either returns its arrays ('Return' in the diagnostic's Features)
or writes the reference's file names into out_folder (joined as a path; the
reference concatenates the folder name and the file name).
"""
from __future__ import annotations

import os

import numpy as np

from .device import DeviceArray, call

PwrFctr = 0.5 * 0.511e6 * 1.6022e-19 / 2.818e-13 * 2.9979e10      # diagnostics.py:24
Ntheta = 60                                                          # diagnostics.py:25


class Diagnostics:
    def __init__(self, Chimera, diags=(), out_folder=None):
        self.Chimera = Chimera
        self.diags = diags
        self.out_folder = out_folder
        if out_folder is not None:
            # created if absent (the reference deletes and recreates it)
            os.makedirs(out_folder, exist_ok=True)

    def do_diags(self, i):
        """diagnostics.py:34-46."""
        for diag in self.diags:
            if np.mod(i, diag["Step"]) != 0:
                continue
            self.istr = str(i).rjust(7, "0")
            diag.setdefault("Features", {})
            kind = diag["Type"]
            if kind == "Fields":
                self.fld_out(diag)
            elif kind == "FieldsFB":
                self.fldfb_out(diag)
            elif kind == "Particles":
                self.phs_out(diag)
            elif kind == "Density":
                self.dns_out(diag)
            elif kind == "EnergyEM":
                self.nrg_out(diag)
            elif kind == "Power":
                self.pwr_out(diag)

    # ------------------------------------------------------------------ dumps (diagnostics.py:50-107)
    def _dump(self, diag, name, dat, ret):
        """Return dat (appended to ret) or save it as out_folder/name, name a
        pattern with {istr} for the zero-padded step of the last do_diags."""
        if "Return" in diag["Features"]:
            ret.append(dat)
            return
        if self.out_folder is None:
            raise ValueError(f"{name} is written into out_folder: give one, or 'Return' in Features")
        np.save(os.path.join(self.out_folder, name.format(istr=getattr(self, "istr", "0000000"))), dat)

    def fld_out(self, diag):
        """diagnostics.py:50-59: EB without its ghost row, per solver, as ee_<n>_<step>.npy."""
        ret = []
        for jj, sol in enumerate(self.Chimera.Solvers):
            self._dump(diag, f"ee_{jj}_{{istr}}.npy", sol.Data["EB"].numpy()[:, 1:], ret)
        if "Return" in diag["Features"]:
            return ret

    def fldfb_out(self, diag):
        """diagnostics.py:61-69: EG_fb per solver (the reference names these
        files ee_<n>_<step>.npy too)."""
        ret = []
        for jj, sol in enumerate(self.Chimera.Solvers):
            self._dump(diag, f"ee_{jj}_{{istr}}.npy", sol.Data["EG_fb"].numpy(), ret)
        if "Return" in diag["Features"]:
            return ret

    def dns_out(self, diag):
        """diagnostics.py:71-88: each pushed species' density on the grid,
        modes 0..MaxMode (default: the first solver's MaxAzimuthMode), ghost
        row dropped, as dens_<n>_<step>.npy."""
        if "MaxMode" in diag["Features"]:
            modes = diag["Features"]["MaxMode"]
        elif len(self.Chimera.Solvers) > 0:
            modes = self.Chimera.Solvers[0].Args["MaxAzimuthMode"]
        else:
            modes = 0
        ret = []
        for jj, sp in enumerate(self.Chimera.Particles):
            if "Still" in sp.Args["Features"]:
                continue
            self._dump(diag, f"dens_{jj}_{{istr}}.npy", sp.get_dens_on_grid(modes)[:, 1:], ret)
        if "Return" in diag["Features"]:
            return ret

    def phs_out(self, diag):
        """diagnostics.py:89-107: half-step positions, momenta and weights of
        each pushed species' particles with sum_c (1 + u_c^2) > 20^2, as a
        (7, n) array in phs<n>_<step>.npy."""
        ret = []
        for jj, sp in enumerate(self.Chimera.Particles):
            if "Still" in sp.Args["Features"]:
                continue
            self._dump(diag, f"phs{jj}_{{istr}}.npy", sp.phase_space(), ret)
        if "Return" in diag["Features"]:
            return ret

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

    def _vec_out(self, sol):
        """fb_vec_out (fb_io.f90:100-140) of the E part of EG_fb with the full
        radial grid on the device, on the transforms' wave numbers (FBoutFull,
        solvers.py:181: kx_env for the envelope solver): the first three components of the returned
        (Nx, Nr_full + 1, M, 6) device array, ghost row zero."""
        A = sol.Args
        Nx, Nkr, M = sol.n_fb
        nr = A["OutFull"].shape[1]
        out_full = self._table(sol, "OutFull", A["OutFull"], sol.rt)
        vec = sol.dev.get("PwrVec")
        if vec is None:
            vec = sol.dev["PwrVec"] = DeviceArray((Nx, nr + 1, M, 6), sol.ct)
        EG = sol.Data["EG_fb"]
        # E in both halves of the transform's input: the first three output components are fb_vec_out
        call(sol.fn("chimera_fb_inverse"), vec.ptr, EG.ptr, EG.ptr, float(A["leftX"]), sol.dev["kx_base"].ptr,
             out_full.ptr, Nx, nr, M, Nkr, 1, None)
        return vec, nr

    def intens_profo(self, sol, NO=None):
        """intens_profO (utils.f90:18-57) of fb_vec_out on the device:
        PWR_RO(NO, nr).  The reference's mode axis is -nkO..nkO, so the
        solver's mode count must be odd (MaxAzimuthMode = 0 for the plain
        solver), as the f2py dimension check requires."""
        NO = Ntheta if NO is None else NO
        vec, nr = self._vec_out(sol)
        Nx, Nkr, M = sol.n_fb
        pwr = np.zeros((NO, nr), order="F")
        call("chimera_intens_profo_c8" if sol.mixed else "chimera_intens_profo", pwr.ctypes.data, vec.ptr, NO, Nx,
             nr, M)
        return pwr

    def pwr_out(self, diag):
        """diagnostics.py:124-149: the field of fb_vec_out (fb_io.f90:100-140;
        the E part of the inverse transform with the full radial grid, ghost
        row zero) reduced on the device to PwrFctr*2*dr * sum_r Rgrid[r]
        sum_m sum_c |E|^2 per x; with 'Spot', also intens_profO of it."""
        ret = []
        spot = "Spot" in diag["Features"]
        for jj, sol in enumerate(self.Chimera.Solvers):
            A = sol.Args
            Nx, Nkr, M = sol.n_fb
            vec, nr = self._vec_out(sol)
            rgrid = self._table(sol, "RgridFull", A["RgridFull"], np.float64)
            dat = np.zeros(Nx)
            call("chimera_line_power_c8" if sol.mixed else "chimera_line_power", vec.ptr, rgrid.ptr, Nx, nr, M,
                 dat.ctypes.data)
            dat = PwrFctr * 2 * A["dr"] * dat
            prof = self.intens_profo(sol) if spot else None
            if "Return" in diag["Features"]:
                ret.append([dat, prof] if spot else dat)
            else:
                if self.out_folder is None:
                    raise ValueError("pwr_out writes pwrX_<n>.txt into out_folder: give one, or 'Return' in Features")
                with open(os.path.join(self.out_folder, f"pwrX_{jj}.txt"), mode="a") as fout:
                    np.savetxt(fout, dat[None])
                if spot:
                    with open(os.path.join(self.out_folder, f"pwrO_{jj}.txt"), mode="a") as fout:
                        np.savetxt(fout, prof.ravel()[None])
        if "Return" in diag["Features"]:
            return ret

    def get_spot_cartesian(self, val=None, solver_index=0, bins=(200, 200), th_part=1.0):
        """diagnostics.py:151-175: the azimuthal intensity profile (intens_profO
        of fb_vec_out, on the device, unless val is given) interpolated from
        the polar (r, theta) nodes onto a bins[0] x bins[1] Cartesian grid of
        half width th_part * max(RgridFull), with scipy's linear griddata as
        the reference; returns (values, extent)."""
        from scipy.interpolate import griddata
        sol = self.Chimera.Solvers[solver_index]
        if val is None:
            val = self.intens_profo(sol, Ntheta)
        r_max = sol.Args["RgridFull"].max()
        rr = np.r_[0:r_max:val.shape[1] * 1j]
        ph = 2 * np.pi / (Ntheta - 1.0) * np.arange(1, Ntheta + 1)
        ph, rr = np.meshgrid(ph, rr)
        coord = ((rr * np.cos(ph)).flatten(), (rr * np.sin(ph)).flatten())
        r_max = th_part * r_max
        new_coord = np.mgrid[-r_max:r_max:bins[0] * 1j, -r_max:r_max:bins[1] * 1j]
        val = griddata(coord, val.T.flatten(), (new_coord[0].flatten(), new_coord[1].flatten()), fill_value=0.0,
                       method="linear").reshape(new_coord[0].shape)
        ext = np.array([-r_max, r_max, -r_max, r_max])
        return val, ext
