"""ChimeraRun: the time loop of the PIC step, driving resident Solver/Specie
objects through libchimera_hip.  Same constructor dictionary and methods as the
reference's ChimeraRun (chimera_main.py:23-318) and the same call order inside
make_halfstep / make_step (chimera_main.py:61-92); the per-particle work of
make_field + proj_fld + make_device + push_velocs runs as one fused gather-push
per species, and push_coords precedes the deposits exactly as in the reference.

SimComps may carry ``'SortEvery': n``: every n steps the pushed species are
sorted by cell on the device (a performance option; the reference sorts only
by x-chunk, chimera_main.py:306-318, and the physics does not depend on the
particle order beyond the order of floating-point sums).  When such a step
moves a window with an absorbing layer, the sort takes the layer's bounds and
drops the absorbed particles itself, in place of the compaction.

SimComps may carry ``'Features': ('PreDepositBackground',)``: the still
background is then deposited before the initial half step, so that the static
initial field of a neutral plasma starts from the neutral charge density (the
reference deposits it only on moving-frame steps, chimera_main.py:277-284, and
its make_halfstep therefore sees electrons without ions).
"""
from __future__ import annotations

import ctypes
import os
import time

import numpy as np

from .device import call


class ChimeraRun:
    def __init__(self, SimComps):
        self.Particles = SimComps.get("Particles", ())
        self.Solvers = SimComps.get("Solvers", ())
        self.MovingFrames = SimComps["MovingFrames"] if "MovingFrames" in SimComps else ({},)
        self.Features = SimComps.get("Features", ())
        self.SortEvery = int(SimComps.get("SortEvery", 0))
        # one precision for the whole run ('fp64' or 'mixed', the Precision key of the components)
        precs = {c.Args.get("Precision", "fp64") for c in tuple(self.Solvers) + tuple(self.Particles)}
        if len(precs) > 1:
            raise ValueError(f"solvers and species must share one Precision, got {sorted(precs)}")
        self.mixed = precs == {"mixed"}
        self._sort_step = False     # this step sorts the pushed species by cell (SortEvery)
        self._sorted = set()        # ids of the species already sorted in this step
        self.init_Moving_Frames()
        self.init_damp_profile()
        if "PreDepositBackground" in self.Features:
            for solver in self.Solvers:
                if "StillAsBackground" in solver.Args["Features"]:
                    self.dep_bg(solver)
        self.make_halfstep()

    def _pfn(self, name):
        """Particle entry point of the run's precision."""
        return name + "_mixed" if self.mixed else name

    def init_Moving_Frames(self):
        for wind in self.MovingFrames:                                     # chimera_main.py:41-52
            wind.setdefault("Features", ())
            wind.setdefault("TimeActive", (0.0, np.inf))
            wind.setdefault("Velocity", 1.0)
            wind.setdefault("TimeStep", 1.0)
            wind.setdefault("Steps", 1)
            f = 0.5 if "Staged" in wind["Features"] else 1.0
            wind["shiftX"] = f * wind["Velocity"] * wind["TimeStep"] * wind["Steps"]

    def init_damp_profile(self):
        for solver in self.Solvers:
            for wind in self.MovingFrames:
                if "AbsorbLayer" in wind:
                    solver.Args["damp_profile"] = solver.get_damp_profile(wind["AbsorbLayer"])

    # ------------------------------------------------------------------ the step
    def make_halfstep(self):
        for species in self.Particles:                                     # chimera_main.py:61-80
            a = self.Solvers[0].Args if len(self.Solvers) > 0 else self.Particles[0].Args
            SimDom = np.asfortranarray([a["leftX"], a["rightX"], 0, a["upperR"] ** 2])
            if SimDom[0] != SimDom[1]:
                species.chunk_and_damp(SimDom=SimDom)
        self.project_current()
        self.project_density()
        for solver in self.Solvers:
            for species in self.Particles:
                solver.maxwell_solver_stat(species.Args["MomentaMeans"][0])
        self.project_fields()
        self.gather_push(0, dt_frac=0.5)

    def make_step(self, istep):
        if self.Solvers and self.Solvers[0].frame is not None:
            call("chimera_frame_step", self.Solvers[0].frame.ptr, *self._frame_shifts(istep))
        self._sort_step = bool(self.SortEvery and np.mod(istep, self.SortEvery) == 0 and self.Solvers)
        self._sorted.clear()
        self.frame_act(istep)                                              # chimera_main.py:82-92
        if self._sort_step:
            for species in self.Particles:
                if "Still" not in species.Args["Features"] and id(species) not in self._sorted:
                    species.sort_cells(self.Solvers[0])
        if self._fusable(istep):
            # push_coords + dep_curr + dep_dens in one pass per species; the
            # transforms then run in the reference's order around stage 2
            self.push_and_deposit(istep)
            for solver in self.Solvers:
                if "NoCurrent" in solver.Args["Features"]:
                    solver.Data["J_fb"].zero_()
                else:
                    solver.fb_curr_in()
            self.frame_act(istep, "stage2")
            for solver in self.Solvers:
                if "StaticKick" in solver.Args["Features"]:
                    # the static density is that of the half-step positions, at the stage-2 edge
                    self.dep_dens(solver, "coords_halfstep")
                    solver.fb_dens_in()
                    solver.FBGradDens()
                elif "SpaceCharge" in solver.Args["Features"]:
                    D = solver.Data
                    D["gradRho_fb_prv"], D["gradRho_fb_nxt"] = D["gradRho_fb_nxt"], D["gradRho_fb_prv"]
                    solver.fb_dens_in()
                    solver.FBGradDens()
        else:
            for species in self.Particles:
                species.push_coords()
            self.chunk_particles(istep)
            self.project_current()
            self.frame_act(istep, "stage2")
            self.project_density()
        self.update_fields()
        self.project_fields()
        self.gather_push(istep)

    def _fusable(self, istep):
        if len(self.Solvers) != 1 or self.Solvers[0].env:
            return False
        for sp in self.Particles:
            if "Xchunked" in sp.Args and "NoSorting" not in sp.Args["Features"] and \
                    np.mod(istep, sp.Args["Xchunked"][1] + 1) == 0:
                return False
        return True

    def _stage2_leftX(self, istep, leftX, Xgrid):
        """leftX after the stage-2 moves of this step, formed as move_frame forms it."""
        for wind in self.MovingFrames:
            if istep < wind["TimeActive"][0] or istep > wind["TimeActive"][1]:
                continue
            if np.mod(istep - wind["TimeActive"][0], wind["Steps"]) != 0:
                continue
            if "Staged" in wind["Features"]:
                Xgrid = Xgrid + wind["shiftX"]
                leftX = Xgrid[0]
        return leftX

    def push_and_deposit(self, istep):
        solver = self.Solvers[0]
        A, D = solver.Args, solver.Data
        curr = None if "NoCurrent" in A["Features"] else D["J"]
        dens = D["Rho"] if "SpaceCharge" in A["Features"] and "StaticKick" not in A["Features"] else None
        if curr is not None:
            curr.zero_()
        if dens is not None:
            if "StillAsBackground" in A["Features"]:
                dens.copy_(D["BckGrndRho"])
            else:
                dens.zero_()
        g = solver.geom_args()
        leftX_rho = self._stage2_leftX(istep, A["leftX"], A["Xgrid"])
        for sp in self.Particles:
            if sp.np == 0 or "Still" in sp.Args["Features"]:
                continue
            call(self._pfn("chimera_soa_push_deposit"), sp.ptr("coords"), sp.ptr("coords_halfstep"), sp.ptr("momenta"),
                 sp.ptr("weights"), sp.ld, sp.np, float(sp.Args["TimeStep"]), g[0], float(leftX_rho), *g[1:],
                 curr.ptr if curr is not None else None, dens.ptr if dens is not None else None, solver.frame_ptr(0))

    # ------------------------------------------------------------------ sources
    def project_current(self):
        for solver in self.Solvers:
            if "NoCurrent" in solver.Args["Features"]:
                solver.Data["J"].zero_()
                solver.Data["J_fb"].zero_()
                continue
            self.dep_curr(solver)
            solver.fb_curr_in()

    def project_density(self):
        for solver in self.Solvers:                                        # chimera_main.py:103-114
            static = "StaticKick" in solver.Args["Features"]
            if "SpaceCharge" not in solver.Args["Features"] and not static:
                continue
            D = solver.Data
            if not static:
                D["gradRho_fb_prv"], D["gradRho_fb_nxt"] = D["gradRho_fb_nxt"], D["gradRho_fb_prv"]
            self.dep_dens(solver, "coords_halfstep" if static else "coords")
            solver.fb_dens_in()
            solver.FBGradDens()

    def dep_curr(self, solver):
        """J = sum of species deposits, unfolded (the fold is applied by fb_curr_in)."""
        J = solver.Data["J"]
        J.zero_()
        g = solver.geom_args()
        for species in self.Particles:
            if species.np == 0 or "Still" in species.Args["Features"]:
                continue
            if solver.env:                                                 # chimera_main.py:158-170
                call("chimera_soa_dep_curr_env", species.ptr("coords_halfstep"), species.ptr("momenta"),
                     species.ptr("weights"), species.ld, species.np, J.ptr, *solver.geom_args_env(),
                     solver.frame_ptr(0))
                continue
            call(self._pfn("chimera_soa_push_deposit"), species.ptr("coords"), species.ptr("coords_halfstep"),
                 species.ptr("momenta"), species.ptr("weights"), species.ld, species.np, 0.0, g[0], g[0], *g[1:],
                 J.ptr, None, None)

    def _dep_dens_env(self, solver, species, component, out):
        call("chimera_soa_dep_dens_env", species.ptr(component), species.ptr("weights"), species.ld, species.np,
             out.ptr, *solver.geom_args_env(), None)

    def dep_dens(self, solver, component="coords"):
        R = solver.Data["Rho"]
        if "StillAsBackground" in solver.Args["Features"]:
            R.copy_(solver.Data["BckGrndRho"])
        else:
            R.zero_()
        g = solver.geom_args()
        for species in self.Particles:
            if species.np == 0 or "Still" in species.Args["Features"]:
                continue
            if solver.env:                                                 # chimera_main.py:195-206
                self._dep_dens_env(solver, species, component, R)
                continue
            call(self._pfn("chimera_soa_push_deposit"), species.ptr(component), species.ptr("coords_halfstep"),
                 species.ptr("momenta"), species.ptr("weights"), species.ld, species.np, 0.0, g[0], g[0], *g[1:],
                 None, R.ptr, None)

    def dep_bg(self, solver):
        B = solver.Data["BckGrndRho"]
        B.zero_()
        g = solver.geom_args()
        for species in self.Particles:
            if species.np == 0 or "Still" not in species.Args["Features"]:
                continue
            if solver.env:                                                 # chimera_main.py:225-236
                self._dep_dens_env(solver, species, "coords", B)
                continue
            call(self._pfn("chimera_soa_push_deposit"), species.ptr("coords"), species.ptr("coords_halfstep"),
                 species.ptr("momenta"), species.ptr("weights"), species.ld, species.np, 0.0, g[0], g[0], *g[1:],
                 None, B.ptr, None)

    # ------------------------------------------------------------------ fields
    def update_fields(self):
        for solver in self.Solvers:                                        # chimera_main.py:116-128
            if "StaticKick" in solver.Args["Features"]:
                solver.Data["EG_fb"].zero_()
                for species in self.Particles:
                    px0 = species.momentum_mean()
                    solver.poiss_corr_stat(px0)
                    solver.maxwell_solver_stat(px0)
                    solver.field_drift(px0)
            else:
                solver.poiss_corr()
                solver.maxwell_solver()

    def project_fields(self):
        for solver in self.Solvers:                                        # chimera_main.py:130-151
            if "NoPush" in solver.Args["Features"]:
                solver.Data["EB"].zero_()
                continue
            solver.G2B_FBRot()
            solver.fb_fld_out()

    def gather_push(self, istep, dt_frac=1.0):
        """make_field + proj_fld (each solver) + make_device + push_velocs, fused
        per species (chimera_main.py:130-151, :91-92; species.py:246-298)."""
        if len(self.Solvers) > 1:
            raise NotImplementedError("one field solver per run is supported")
        for species in self.Particles:
            if species.np == 0 or "Still" in species.Args["Features"]:
                continue
            A = species.Args
            devs = species.device_descriptors(istep)
            push_dt = float(A["push_fact"] * dt_frac * A["TimeStep"])
            frame = None
            if self.Solvers and "NoPush" not in self.Solvers[0].Args["Features"]:
                solver = self.Solvers[0]
                EB, g, frame = solver.Data["EB"].ptr, solver.geom_args(), solver.frame_ptr(1)
            else:
                EB, g = 0, (0.0, 0, 1.0, 1.0, 0.0, -1.0, 1, 1, 0)   # rmax < 0: nothing gathered
            recs, t = devs if devs is not None else (None, 0.0)
            if self.Solvers and self.Solvers[0].env and EB:                # proj_fld_env, chimera_main.py:142-146
                call("chimera_soa_gather_push_env", species.ptr("coords"), species.ptr("momenta"),
                     species.ptr("weights"), species.ld, species.np, EB, *self.Solvers[0].geom_args_env(), push_dt,
                     ctypes.addressof(recs) if recs is not None else None, len(recs) if recs is not None else 0, t,
                     frame)
                continue
            call(self._pfn("chimera_soa_gather_push_dev"), species.ptr("coords"), species.ptr("momenta"),
                 species.ptr("weights"), species.ld, species.np, EB, *g, push_dt,
                 ctypes.addressof(recs) if recs is not None else None, len(recs) if recs is not None else 0, t, frame)

    # ------------------------------------------------------------------ moving frames (chimera_main.py:250-318)
    def damp_fields(self, wind):
        if "AbsorbLayer" not in wind or wind["AbsorbLayer"] <= 0:
            return
        for solver in self.Solvers:
            if "StaticKick" not in solver.Args["Features"]:
                solver.damp_field()

    def damp_plasma(self, wind):
        """The absorbing layer on the particles (species.py:351-398).  On a
        step that sorts by cell, a pushed species is sorted here instead, with
        the layer's bounds: the sort drops the same particles, so the one
        pass replaces the compaction."""
        if "AbsorbLayer" not in wind:
            return
        for species in self.Particles:
            if self._sort_step and "Still" not in species.Args["Features"]:
                species.sort_cells(self.Solvers[0], SimDom=species.damp_lims(wind))
                self._sorted.add(id(species))
            else:
                species.chunk_and_damp(wind=wind)

    def add_plasma(self, wind):
        if "AddPlasma" not in wind:
            return
        xs = int(wind["shiftX"] / wind["TimeStep"]) + 1
        if "IonsOnTop" in wind["Features"]:
            s1, s2 = self.Particles[:2]
            parts = s1.gen_parts(Xsteps=xs, ProfileFunc=wind["AddPlasma"])
            s1.add_particles(*parts)
            parts[-1] = parts[-1] * -1
            s2.add_particles(*parts)
        else:
            for species in self.Particles:
                species.add_particles(*species.gen_parts(Xsteps=xs, ProfileFunc=wind["AddPlasma"]))

    def postframe_corr(self, wind):
        # chimera_main.py:277-284: the condition is always true in the reference
        for solver in self.Solvers:
            if "StaticKick" in solver.Args["Features"]:
                continue
            if "SpaceCharge" in solver.Args["Features"]:
                if "StillAsBackground" in solver.Args["Features"]:
                    self.dep_bg(solver)
                # the reference also re-deposits Rho here; project_density overwrites it before any use

    def move_frame(self, wind):
        for comp in tuple(self.Solvers) + tuple(self.Particles):
            comp.Args["Xgrid"] = comp.Args["Xgrid"] + wind["shiftX"]
            comp.Args["leftX"] = comp.Args["Xgrid"][0]
            comp.Args["rightX"] = comp.Args["Xgrid"][-1]

    def frame_act(self, istep, act="stage1"):
        for wind in self.MovingFrames:
            if istep < wind["TimeActive"][0] or istep > wind["TimeActive"][1]:
                continue
            if np.mod(istep - wind["TimeActive"][0], wind["Steps"]) != 0:
                continue
            if act == "stage1":
                self.damp_fields(wind)
                self.move_frame(wind)
                self.add_plasma(wind)
                self.damp_plasma(wind)
                self.postframe_corr(wind)
            elif act == "stage2" and "Staged" in wind["Features"]:
                self.move_frame(wind)

    # ------------------------------------------------------------------ device frame and captured steps
    def graph_blockers(self):
        """Why a step of this run cannot be recorded once and replayed (an
        empty list when it can): every step must launch the same work on the
        same arrays, so nothing may add, absorb or re-sort particles, and the
        window may move only by the same amount every step."""
        out = []
        if len(self.Solvers) != 1:
            out.append("needs exactly one field solver")
        if self.SortEvery:
            out.append("SortEvery re-sorts on some steps")
        for sp in self.Particles:
            if "Xchunked" in sp.Args and "NoSorting" not in sp.Args["Features"]:
                out.append("Xchunked sorting runs on some steps")
        if len(self.MovingFrames) > 1:
            out.append("more than one moving frame")
        for wind in self.MovingFrames:
            if wind["Steps"] != 1 or wind["TimeActive"][0] > 1 or wind["TimeActive"][1] != np.inf:
                out.append("the window does not move on every step")
            if "AddPlasma" in wind or "AbsorbLayer" in wind:
                out.append("plasma injection or absorption changes the particle count")
            if self.Solvers and "StillAsBackground" in self.Solvers[0].Args["Features"]:
                out.append("the still background is re-deposited after each move")
        if self.Solvers and "StaticKick" in self.Solvers[0].Args["Features"]:
            out.append("StaticKick reads the mean momentum back to the host on every step")
        if self.Solvers and self.Solvers[0].env:
            out.append("the envelope (KxShift) solver runs the unfused particle calls")
        for sp in self.Particles:
            for dev in getattr(sp, "Devices", ()):
                fn = dev[0]
                name = (fn if isinstance(fn, str) else getattr(fn, "__name__", "")).lower()
                if name in ("planewave", "gaussbeam"):
                    out.append(f"the {name} device depends on the step's time")
        return out

    def enable_device_frame(self):
        """Keep the window edge in a device frame that a kernel advances at the
        start of every step (frame[0]: after the stage-1 move, frame[1]: after
        the stage-2 move), so that a recorded step replays with the moving
        window.  The host Args keep moving exactly as before."""
        from .device import DeviceArray
        blockers = self.graph_blockers()
        if blockers:
            raise ValueError("device frame not possible: " + "; ".join(blockers))
        s = self.Solvers[0]
        if s.frame is None:
            x = float(s.Args["leftX"])
            s.frame = DeviceArray.from_numpy(np.array([x, x]), np.float64)

    def _frame_shifts(self, istep):
        """(shift at stage 1, shift at stage 2) of step istep, as frame_act moves the window."""
        s1 = s2 = 0.0
        for wind in self.MovingFrames:
            if istep < wind["TimeActive"][0] or istep > wind["TimeActive"][1]:
                continue
            if np.mod(istep - wind["TimeActive"][0], wind["Steps"]) != 0:
                continue
            s1 = float(wind["shiftX"])
            if "Staged" in wind["Features"]:
                s2 = float(wind["shiftX"])
        return s1, s2

    def host_step(self, istep):
        """The host-side effects of make_step(istep) -- window moves and the
        swap of the two density-gradient buffers -- without launching anything:
        the device work of the step is replayed from a recorded graph."""
        for wind in self.MovingFrames:
            if istep < wind["TimeActive"][0] or istep > wind["TimeActive"][1]:
                continue
            if np.mod(istep - wind["TimeActive"][0], wind["Steps"]) != 0:
                continue
            self.move_frame(wind)
        for solver in self.Solvers:
            if "SpaceCharge" in solver.Args["Features"]:
                D = solver.Data
                D["gradRho_fb_prv"], D["gradRho_fb_nxt"] = D["gradRho_fb_nxt"], D["gradRho_fb_prv"]
        for wind in self.MovingFrames:
            if istep < wind["TimeActive"][0] or istep > wind["TimeActive"][1]:
                continue
            if np.mod(istep - wind["TimeActive"][0], wind["Steps"]) != 0:
                continue
            if "Staged" in wind["Features"]:
                self.move_frame(wind)

    def chunk_particles(self, istep=0):
        for species in self.Particles:                                     # chimera_main.py:306-318
            if "Xchunked" not in species.Args or "NoSorting" in species.Args["Features"]:
                continue
            if np.mod(istep, species.Args["Xchunked"][1] + 1) != 0:
                continue
            if species.np == 0:
                continue
            a = self.Solvers[0].Args if len(self.Solvers) > 0 else self.Particles[0].Args
            SimDom = np.asfortranarray([a["leftX"], a["rightX"], 0, a["Rgrid"].max() ** 2])
            species.chunk_and_damp(SimDom=SimDom, position="cntr")

    # ------------------------------------------------------------------ snapshots (chimera_main.py:323-410)
    @staticmethod
    def _snap_value(v):
        """The array a snapshot stores for one Data/Args value, or None for the
        values the reference skips (containers) and those an archive cannot hold
        without pickling (functions, objects)."""
        from .device import DeviceArray
        if isinstance(v, DeviceArray):
            return v.numpy()
        if isinstance(v, (list, tuple, set, dict)) or callable(v) or v is None:
            return None
        a = np.array(v)                       # a copy: the snapshot must not follow the run
        return None if a.dtype == object else a

    def drop_snap(self, fname="./snap_", verbose=False):
        """chimera_main.py:323-366: every solver's and species' Data arrays and
        plain Args under the reference's names ('solver<i>/Data/<key>',
        'specie<i>/Args/<key>', 'NumSolvers', 'NumSpecies', 'NumThreads'),
        copied from the device.  The container is a NumPy .npz archive (h5py,
        which the reference writes with, is not a dependency here); species
        arrays are stored in the reference's layout, (3, np) and (np,), so only
        live particles are kept.  Returns the file name."""
        fname += time.ctime().replace(" ", "_").replace(":", "-") + ".npz"
        np.savez(fname, **self.snapshot(verbose))
        return fname

    def snapshot(self, verbose=False):
        """The arrays of drop_snap, in host memory (name -> array)."""
        out = {}
        for i, solver in enumerate(self.Solvers):
            for grp, d in (("Data", solver.Data), ("Args", solver.Args)):
                for key, v in d.items():
                    a = self._snap_value(v)
                    if a is not None:
                        if verbose:
                            print(f"solver{i}/{grp}/{key}")
                        out[f"solver{i}/{grp}/{key}"] = a
        out["NumSolvers"] = np.array(len(self.Solvers))
        for i, sp in enumerate(self.Particles):
            for key in sp.Data:
                out[f"specie{i}/Data/{key}"] = sp.host(key)
            for key, v in sp.Args.items():
                a = self._snap_value(v)
                if a is not None:
                    out[f"specie{i}/Args/{key}"] = a
        out["NumSpecies"] = np.array(len(self.Particles))
        out["NumThreads"] = np.array(int(os.environ.get("OMP_NUM_THREADS", "1")))
        for i, sp in enumerate(self.Particles):
            # a species' own random stream (own_rng; the plasma a moving window
            # injects is drawn from it): its state, so that a resumed run draws
            # what the uninterrupted run would.  The global NumPy stream, which
            # the reference draws from, is not the run's to save.
            if isinstance(sp.rng, np.random.RandomState):
                name, keys, pos, has_gauss, gauss = sp.rng.get_state()
                out[f"specie{i}/rng/keys"] = np.asarray(keys, dtype=np.uint32)
                out[f"specie{i}/rng/state"] = np.array([pos, has_gauss], dtype=np.int64)
                out[f"specie{i}/rng/gauss"] = np.array(gauss, dtype=np.float64)
        return out

    def read_snap(self, fname="./test.npz", verbose=False):
        """chimera_main.py:368-410: load a snapshot of drop_snap (a file name, or
        the dict of snapshot()) into this run's solvers and species (same
        configuration), key by key of their own Data and Args; device arrays are
        overwritten in place, so recorded steps keep their buffers.  A device
        window edge restarts from the restored leftX."""
        if isinstance(fname, dict):
            self._load_snap(fname, verbose)
        else:
            with np.load(fname, allow_pickle=False) as f:
                self._load_snap(f, verbose)

    def _load_snap(self, f, verbose):
        """Load from a mapping of snapshot names to arrays (an npz file or a dict)."""
        from .device import DeviceArray
        for i in range(int(f["NumSolvers"])):
            solver = self.Solvers[i]
            for key, v in list(solver.Data.items()):
                k = f"solver{i}/Data/{key}"
                if self._snap_value(v) is None or k not in f:
                    continue
                if verbose:
                    print(k)
                if isinstance(v, DeviceArray):
                    v.upload(f[k])
                else:
                    solver.Data[key] = f[k]
            for key, v in list(solver.Args.items()):
                k = f"solver{i}/Args/{key}"
                if self._snap_value(v) is not None and k in f:
                    a = f[k]
                    solver.Args[key] = a[()] if a.ndim == 0 else a
            if solver.frame is not None:
                x = float(solver.Args["leftX"])
                solver.frame.upload(np.array([x, x]))
        for i in range(int(f["NumSpecies"])):
            sp = self.Particles[i]
            arrays = {key: f[f"specie{i}/Data/{key}"] for key in ("coords", "coords_halfstep", "momenta", "weights")}
            sp.restore_particles(**arrays)
            for key, v in list(sp.Args.items()):
                k = f"specie{i}/Args/{key}"
                if self._snap_value(v) is not None and k in f:
                    a = f[k]
                    sp.Args[key] = a[()] if a.ndim == 0 else a
            if f"specie{i}/rng/keys" in f and isinstance(sp.rng, np.random.RandomState):
                pos, has_gauss = (int(v) for v in f[f"specie{i}/rng/state"])
                sp.rng.set_state(("MT19937", np.asarray(f[f"specie{i}/rng/keys"], dtype=np.uint32), pos, has_gauss,
                                  float(f[f"specie{i}/rng/gauss"])))
