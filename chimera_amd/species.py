"""Specie: a particle species with its particles resident in device memory.

Same constructor dictionary, attributes and methods as the reference's Specie
(species.py:24-398).  Particles are kept structure-of-arrays on the device:
``coords``, ``coords_halfstep``, ``momenta`` as (3, capacity) and ``weights`` as
(capacity,), of which the first ``np`` columns are live.  gen_parts returns host
arrays exactly as the reference does (the random draws are NumPy's legacy
generator in the reference's order, the FixedCell lattice is built by
libchimera_hip); add_particles uploads them.

Extra key 'Precision': 'fp64' (default) or 'mixed' (momenta and weights
float32 on the device; positions stay FP64).
"""
from __future__ import annotations

import ctypes
import inspect

import numpy as np

from . import fimera as F
from .device import DeviceArray, call, copy_elems


class DeviceRecord(ctypes.Structure):
    """chimera_device_t of include/chimera_hip.h."""
    _fields_ = [("kind", ctypes.c_int32), ("nmap", ctypes.c_int32), ("map", ctypes.c_void_p), ("a0", ctypes.c_double),
                ("params", ctypes.c_double * 8)]


class Specie:
    def __init__(self, specie_in):
        self.Args = specie_in
        A = self.Args
        A.setdefault("Features", ())
        A.setdefault("Charge", -1.0)
        A.setdefault("Mass", 1.0)
        A.setdefault("Density", 0.0)
        A.setdefault("MomentaMeans", (0.0, 0.0, 0.0))
        A.setdefault("MomentaSpreads", (0.0, 0.0, 0.0))
        A.setdefault("Grid", (0.0, 0.0, 0.0, 1.0, 1.0))
        self.Devices = A["Devices"] if "Devices" in A else ()
        A.setdefault("Precision", "fp64")
        if A["Precision"] not in ("fp64", "mixed"):
            raise ValueError(f"Precision {A['Precision']!r}: 'fp64' or 'mixed'")
        self.mixed = A["Precision"] == "mixed"
        self.rt = np.float32 if self.mixed else np.float64

        A["leftX"], A["rightX"], A["lengthR"], A["dx"], A["dr"] = A["Grid"]
        A["push_fact"] = 2 * np.pi * A["Charge"] / A["Mass"]             # species.py:64
        if "Xchunked" in A:
            A["nthrds"] = int(A["Xchunked"][0])
            A["Nx"] = 2 * A["nthrds"] * int(np.round(0.5 / A["dx"] / A["nthrds"] * (A["rightX"] - A["leftX"])))
        else:
            A["Nx"] = 2 * int(np.round(0.5 / A["dx"] * (A["rightX"] - A["leftX"])))
        if A["Nx"] > 0:
            A["Xgrid"] = A["rightX"] - A["dx"] * np.arange(A["Nx"])[::-1]
        else:
            A["Xgrid"] = np.array([0.0])
        A["leftX"] = A["Xgrid"][0]
        A["Nr"] = int(np.round(A["lengthR"] / A["dr"]))
        A["Rgrid"] = A["dr"] * (np.arange(A["Nr"]) - 0.5) if A["Nr"] > 0 else np.array([0.0])
        A["lowerR"] = (A["Rgrid"] * (A["Rgrid"] >= 0)).min()
        A["upperR"] = A["Rgrid"].max()

        if "FixedCell" in A:                                                 # species.py:95-107
            A["Num_p"] = int(np.prod(A["FixedCell"]))
            fc = A["FixedCell"]
            packX, packR, packO = np.mgrid[1:fc[0]:fc[0] * 1j, 1:fc[1]:fc[1] * 1j, 1:fc[2]:fc[2] * 1j]
            packX = np.asfortranarray((packX.ravel() - 0.5) / fc[0])
            packR = np.asfortranarray((packR.ravel() - 0.5) / fc[1])
            packO = np.asfortranarray(np.exp(2.0j * np.pi * (packO.ravel() - 1) / fc[2]))
            A["Pax"] = (packX, packR, packO)
        elif "RandCell" in A:                                                # species.py:108-109
            A["Num_p"] = int(A["RandCell"])
        else:
            A["Num_p"] = 0
        A["wght0"] = (A["Charge"] * A["Density"] * A["dr"] * A["dx"] * 2 * np.pi / A["Num_p"]) if A["Num_p"] else 0.0
        A["NpSlice"] = A["Nx"] * A["Num_p"]

        # random draws of gen_parts: NumPy's global legacy stream, as in the
        # reference; a simulation that must not share it (several runs in one
        # process) sets its own numpy.random.RandomState here
        self.rng = np.random
        self.np = 0
        self.cap = 0
        self.Data = {}
        self._alloc(1024)

    # ------------------------------------------------------------------ storage
    # coords, coords_halfstep, momenta: component c of particle i at c*cap + i
    # (a Fortran (cap, 3) array); weights: (cap,).
    def _alloc(self, cap):
        cap = -(-int(cap) // 64) * 64        # components start on 512-byte boundaries (vector loads)
        old, oldcap = (self.Data, self.cap) if self.cap else (None, 0)
        D = {k: DeviceArray((cap, 3), np.float64) for k in ("coords", "coords_halfstep")}
        D["momenta"] = DeviceArray((cap, 3), self.rt)
        D["weights"] = DeviceArray((cap,), self.rt)
        if old is not None and self.np:
            for k in ("coords", "coords_halfstep", "momenta"):
                for c in range(3):
                    copy_elems(D[k], c * cap, old[k], c * oldcap, self.np)
            copy_elems(D["weights"], 0, old["weights"], 0, self.np)
        self.Data = D
        self.cap = cap

    @property
    def ld(self):
        return self.cap

    def ptr(self, key):
        return self.Data[key].ptr

    def host(self, key):
        """Live part of a particle array as a host array in the reference's layout."""
        a = self.Data[key].numpy()
        if key == "weights":
            return a[:self.np].copy()
        return np.asfortranarray(a[:self.np].T)

    # ------------------------------------------------------------------ diagnostics
    def get_dens_on_grid(self, Nko=0):
        """species.py:400-427: the species' density on its grid, modes 0..Nko,
        complex (Nx, Nr, Nko+1): dep_dens of the particles times
        1/(2 pi dx dr Rgrid) per radial row (0 at Rgrid <= 0), deposited on
        the device (FP64 sums in every precision) and returned to the host."""
        A = self.Args
        R = A["Rgrid"]
        V = 2 * np.pi * A["dx"] * A["dr"] * R
        V = (V + (R == 0)) ** -1 * (R > 0.0)
        if getattr(self, "_dens_tabs", (None,))[0] is not R:
            self._dens_tabs = (R, DeviceArray.from_numpy(np.asarray(R, float), np.float64),
                               DeviceArray.from_numpy(np.asarray(V, float), np.float64))
        _, Rdev, Vdev = self._dens_tabs
        dens = DeviceArray((A["Nx"], A["Nr"], Nko + 1), np.complex128)
        call("chimera_soa_dens_on_grid_mixed" if self.mixed else "chimera_soa_dens_on_grid", self.ptr("coords"),
             self.ptr("weights"), self.ld, self.np, dens.ptr, float(A["leftX"]), Rdev.ptr, 1.0 / A["dx"],
             1.0 / A["dr"], float(R[0]), float(R[-1]), A["Nx"] - 1, A["Nr"] - 1, int(Nko), Vdev.ptr)
        return dens.numpy()

    def phase_space(self):
        """The PartsPack of phs_out (diagnostics.py:94-103): half-step
        positions, momenta and weights of the particles with
        sum_c (1 + u_c^2) > 20^2, as a (7, n) array, selected on the device."""
        fn = "chimera_phs_select_mixed" if self.mixed else "chimera_phs_select"
        n = ctypes.c_int64(0)
        args = (self.ptr("coords_halfstep"), self.ptr("momenta"), self.ptr("weights"), self.ld, self.np)
        call(fn, *args, None, 0, ctypes.addressof(n))
        if n.value == 0:
            return np.zeros((7, 0), order="F")
        out = DeviceArray((7, n.value), np.float64)
        call(fn, *args, out.ptr, n.value, ctypes.addressof(n))
        return out.numpy()

    # ------------------------------------------------------------------ generation (species.py:132-244)
    def gen_parts(self, Domain=None, Xsteps=None, ProfileFunc=None):
        A = self.Args
        Xgrid, Rgrid = A["Xgrid"], A["Rgrid"]
        if Domain is not None:
            p_left, p_right, p_low, p_up = Domain
            if A["leftX"] > p_right or A["rightX"] < p_left or p_up < A["lowerR"] or p_low > A["upperR"]:
                return [np.zeros((3, 0), order="F"), np.zeros((3, 0), order="F"), np.zeros(0)]
            ixb, ixe = (Xgrid < p_left).sum() - 1, (Xgrid < p_right).sum() + 1
            Xgrid = Xgrid[ixb:ixe]
            irb = 0 if p_low <= Rgrid.min() else (Rgrid < p_low).sum() - 1
            ire = Rgrid.shape[0] if p_up >= Rgrid.max() else (Rgrid < p_up).sum() + 1
            Rgrid = Rgrid[irb:ire]
        elif Xsteps is not None:
            Xgrid = Xgrid[-2 * Xsteps:-Xsteps]
        if A["Num_p"] == 0:
            return [np.zeros((3, 0), order="F"), np.zeros((3, 0), order="F"), np.zeros(0)]
        coords = np.zeros((4, Xgrid.shape[0] * Rgrid.shape[0] * A["Num_p"]), order="F")
        if "FixedCell" in A:
            RandPackO = self.rng.rand(Xgrid.shape[0], Rgrid.shape[0]).astype("d", order="F")
            coords, n = F.genparts(coords, Xgrid, Rgrid, RandPackO, *A["Pax"])
        else:
            coords, n = self.gen_randcell(coords, Xgrid, Rgrid)
        coords = coords[:, :n]
        if ProfileFunc is None:
            coords[-1] *= A["wght0"]
        else:
            nargs = len(inspect.signature(ProfileFunc).parameters)
            if nargs == 1:
                coords[-1] *= A["wght0"] * ProfileFunc(coords[0])
            elif nargs == 2:
                coords[-1] *= A["wght0"] * ProfileFunc(coords[0], np.sqrt((coords[1:3] ** 2).sum(0)))
            elif nargs == 3:
                coords[-1] *= A["wght0"] * ProfileFunc(*coords[0:3])
            else:
                raise ValueError("can't understand the ProfileFunc")
        if "FlatSpectrum" in A["Features"]:
            rand_mom = 2 * self.rng.rand(3, n) - 1.0
        else:
            rand_mom = self.rng.randn(3, n)
        mm, ms = A["MomentaMeans"], A["MomentaSpreads"]
        momenta = np.vstack([mm[c] + ms[c] * rand_mom[c] for c in range(3)]).astype("d", order="F")
        weights = coords[-1].astype("d", order="F")
        return [np.asfortranarray(coords[0:3]), momenta, weights]

    def gen_randcell(self, coords, Xgrid, Rgrid):
        """species.py:429-471: Num_p particles in each cell between neighbouring
        nodes, stratified per axis (the k-th of Num_p sub-intervals, jittered
        by up to half its width) and paired across axes by three shuffles.
        A set-up step on the host: the legacy generator's shuffle draws a
        data-dependent number of values, so the cells are drawn one by one,
        in the reference's order, to keep its random stream."""
        npc, dx, dr = self.Args["Num_p"], self.Args["dx"], self.Args["dr"]
        k = np.arange(npc, dtype=np.float64)
        ip = 0
        for ix in range(Xgrid.shape[0] - 1):
            for ir in range(Rgrid.shape[0] - 1):
                u = 2 * (self.rng.rand(3, npc) - 0.5)
                x = Xgrid[ix] + dx * (k + 0.5 * u[0]) / npc
                r = Rgrid[ir] + dr * (k + 0.5 * u[1]) / npc
                o = 2 * np.pi * (k + 0.5 * u[2]) / npc
                self.rng.shuffle(x)
                self.rng.shuffle(r)
                self.rng.shuffle(o)
                sl = slice(ip, ip + npc)
                coords[0, sl], coords[1, sl], coords[2, sl], coords[3, sl] = x, r * np.cos(o), r * np.sin(o), r
                ip += npc
        return coords, ip

    def add_particles(self, coords, momenta, weights):
        """Append host particles (species.py:217-244); the half-step copy starts at coords."""
        n = coords.shape[-1]
        if n == 0:
            return
        if self.np + n > self.cap:
            self._alloc(int((self.np + n) * 1.25) + 1024)
        D, cap, o = self.Data, self.cap, self.np
        for c in range(3):
            D["coords"].write(c * cap + o, coords[c])
            D["coords_halfstep"].write(c * cap + o, coords[c])
            D["momenta"].write(c * cap + o, momenta[c])
        D["weights"].write(o, weights)
        self.np += n

    def restore_particles(self, coords, coords_halfstep, momenta, weights):
        """Replace the live particles by host arrays in the reference's layout
        ((3, np) and (np,)), the half-step positions included (read_snap)."""
        n = weights.shape[-1]
        if n > self.cap:
            self.np = 0
            self._alloc(int(n * 1.25) + 1024)
        D, cap = self.Data, self.cap
        for c in range(3):
            D["coords"].write(c * cap, coords[c])
            D["coords_halfstep"].write(c * cap, coords_halfstep[c])
            D["momenta"].write(c * cap, momenta[c])
        D["weights"].write(0, weights)
        self.np = n

    def momentum_mean(self):
        """sum(momenta[0] * weights) / sum(weights) on the device (the PXmean of
        the StaticKick field update, chimera_main.py:120-121); NaN without
        particles, as the reference's quotient gives."""
        out = np.zeros(1)
        call("chimera_momentum_mean_mixed" if self.mixed else "chimera_momentum_mean", self.ptr("momenta"),
             self.ptr("weights"), self.np, out.ctypes.data)
        return float(out[0])

    def beam_focus(self, x_foc):
        """species.py:473-486: move a relativistic bunch back along its ballistic
        drift so that it focuses x_foc downstream (a set-up step, done on host
        copies with the reference's NumPy arithmetic and written back)."""
        coords, momenta = self.host("coords"), self.host("momenta").astype(np.float64)
        weights = self.host("weights").astype(np.float64)
        gg = (1 + momenta ** 2).sum(0) ** 0.5
        pzmean = (momenta[0] * weights).sum() / weights.sum()
        coords[0] = coords[0] - x_foc * (momenta[0] - pzmean) / pzmean / gg ** 2
        coords[1:3] = coords[1:3] - momenta[1:3] / pzmean * x_foc
        self.restore_particles(coords, coords, self.host("momenta"), self.host("weights"))

    def denoise(self, WaveNums2Kill):
        """Mirror particles in anti-phase (species.py:322-349)."""
        if self.mixed:
            raise NotImplementedError("denoise is implemented for FP64 species")
        for k_supp in WaveNums2Kill:
            n = self.np
            if 2 * n > self.cap:
                self._alloc(2 * n + 1024)
            D, cap = self.Data, self.cap
            for key in ("coords", "coords_halfstep", "momenta"):
                for c in range(3):
                    copy_elems(D[key], c * cap + n, D[key], c * cap, n)
            for key in ("coords", "coords_halfstep"):
                call("chimera_axpb", D[key].offset_ptr(n), n, 1.0, 0.5 / k_supp)
            copy_elems(D["weights"], n, D["weights"], 0, n)
            call("chimera_axpb", D["weights"].ptr, 2 * n, 0.5, 0.0)
            self.np = 2 * n

    # ------------------------------------------------------------------ absorbing boundary (species.py:351-398)
    def chunk_and_damp(self, wind=None, SimDom=None, position="stag"):
        """Drop the particles outside SimDom, keeping the order of the others.
        The reference tests coords_halfstep only on the chunked path with
        position='cntr' (species.py:378-386); so does this."""
        if self.np == 0:
            return
        A = self.Args
        if wind is None:
            wind = {"Features": (), "AbsorbLayer": 0.0}
        if SimDom is None:
            SimDom = self.damp_lims(wind)
        half = ("Xchunked" in A and "NoSorting" not in wind["Features"] and position == "cntr")
        lims = np.ascontiguousarray(SimDom, dtype=np.float64)
        nout = ctypes.c_int32(0)
        # the survivors go to a spare set of arrays, which then becomes the
        # species' arrays: no copy back (the spares stay allocated)
        spare = getattr(self, "_spare", None)        # shared with sort_cells
        if spare is None or spare["weights"].shape[0] != self.cap:
            spare = {k: DeviceArray((self.cap, 3), np.float64) for k in ("coords", "coords_halfstep")}
            spare["momenta"] = DeviceArray((self.cap, 3), self.rt)
            spare["weights"] = DeviceArray((self.cap,), self.rt)
        call("chimera_soa_compact_into_mixed" if self.mixed else "chimera_soa_compact_into", self.ptr("coords"),
             self.ptr("coords_halfstep"), self.ptr("momenta"), self.ptr("weights"), self.ld, self.np,
             spare["coords"].ptr, spare["coords_halfstep"].ptr, spare["momenta"].ptr, spare["weights"].ptr, self.cap,
             lims.ctypes.data, 1 if half else 0, ctypes.addressof(nout))
        self._spare, self.Data = self.Data, spare
        self.np = int(nout.value)

    def damp_lims(self, wind):
        """The bounds kept by the absorbing layer of `wind` (species.py:360-362):
        x from the layer's inner edge to the right edge, r^2 up to upperR^2."""
        A = self.Args
        return np.array([A["leftX"] + wind["AbsorbLayer"] * A["dx"], A["rightX"], 0.0, A["upperR"] ** 2])

    def sort_cells(self, solver, SimDom=None, position="stag"):
        """Stable counting sort of the particles by cell of `solver`'s grid,
        dropping those outside SimDom when it is given (libchimera_hip
        chimera_soa_sort_cells).  Keeps the physics and changes the order, so
        that deposition and gather touch neighbouring nodes."""
        if self.np == 0:
            return
        spare = getattr(self, "_spare", None)        # shared with chunk_and_damp
        if spare is None or spare["weights"].shape[0] != self.cap:
            spare = {k: DeviceArray((self.cap, 3), np.float64) for k in ("coords", "coords_halfstep")}
            spare["momenta"] = DeviceArray((self.cap, 3), self.rt)
            spare["weights"] = DeviceArray((self.cap,), self.rt)
        g = solver.geom_args()
        lims = np.ascontiguousarray(SimDom, dtype=np.float64) if SimDom is not None else None
        nout = ctypes.c_int32(0)
        call("chimera_soa_sort_cells_mixed" if self.mixed else "chimera_soa_sort_cells", self.ptr("coords"), self.ptr("coords_halfstep"), self.ptr("momenta"),
             self.ptr("weights"), self.ld, self.np, spare["coords"].ptr, spare["coords_halfstep"].ptr,
             spare["momenta"].ptr, spare["weights"].ptr, self.cap, g[0], g[2], g[3], g[4], g[6], g[7],
             lims.ctypes.data if lims is not None else None, 1 if position == "cntr" else 0, ctypes.addressof(nout),
             None)
        self._spare, self.Data = self.Data, spare
        self.np = int(nout.value)

    # ------------------------------------------------------------------ pushes
    def push_coords(self, dt_frac=1.0):
        if self.np == 0 or "Still" in self.Args["Features"]:
            return
        if self.mixed:
            raise NotImplementedError("mixed-precision species push positions inside the fused deposition")
        call("chimera_soa_push_coords", self.ptr("coords"), self.ptr("coords_halfstep"), self.ptr("momenta"), self.ld,
             self.np, float(dt_frac * self.Args["TimeStep"]))

    # the devices of devices.f90 by their f2py names: (kind, parameter count, has a field map, has an amplitude)
    DEVICE_KINDS = {"undul_analytic": (0, 4, False, False), "undul_analytic_taper": (1, 5, False, False),
                    "undul_mapped": (2, 3, True, False), "undul_mapped_tap": (3, 5, True, False),
                    "planewave": (4, 7, False, False), "gaussbeam": (5, 8, False, True)}

    def device_descriptors(self, i_step):
        """The species' Devices (species.py:258-277: entries (device, *args),
        the device an f2py name or a function of that name, args as the
        reference passes them) as chimera_device_t records for the fused
        gather-push at time i_step * TimeStep, or None without devices.  Field
        maps are uploaded once and kept with the species."""
        if not self.Devices:
            return None
        if len(self.Devices) > 4:
            raise NotImplementedError("at most 4 devices per species are fused into the push")
        recs = (DeviceRecord * len(self.Devices))()
        keep = getattr(self, "_dev_maps", None)
        if keep is None:
            keep = self._dev_maps = {}
        for k, dev in enumerate(self.Devices):
            fn = dev[0]
            name = (fn if isinstance(fn, str) else getattr(fn, "__name__", "")).lower()
            if name not in self.DEVICE_KINDS:
                raise NotImplementedError(f"device {name!r}: the fused push provides {sorted(self.DEVICE_KINDS)}")
            kind, npar, has_map, has_amp = self.DEVICE_KINDS[name]
            args = list(dev[1:])
            rec = recs[k]
            rec.kind = kind
            if has_map:
                src = args.pop(0)
                a0 = np.asarray(src, dtype=np.float64)
                if a0.ndim != 2 or a0.shape[0] != 2:
                    raise ValueError(f"{name}: the map a0 has shape (2, nx), not {a0.shape}")
                if k not in keep or keep[k][0] is not src:
                    keep[k] = (src, DeviceArray.from_numpy(np.asfortranarray(a0)))
                rec.map = keep[k][1].ptr
                rec.nmap = a0.shape[1]
            if has_amp:
                rec.a0 = float(args.pop(0))
            params = np.asarray(args[0], dtype=np.float64).ravel()
            if params.size < npar:
                raise ValueError(f"{name}: {params.size} parameters, {npar} needed")
            for i in range(npar):
                rec.params[i] = float(params[i])
        return recs, i_step * self.Args["TimeStep"]
