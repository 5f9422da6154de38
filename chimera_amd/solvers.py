"""Solver: the quasi-cylindrical PSATD field solver with its state resident in
device memory.  Same constructor dictionary, attribute names and method names as
the reference's Solver (solvers.py:26-225); the field arrays in ``Data`` are
DeviceArrays of the reference's shapes, and every method runs libchimera_hip.

Supported: the plain solver with or without SpaceCharge and
StillAsBackground, Rcut, NoPoissonCorrection, StaticKick, Xchunked grids,
CoPropagative; and the envelope (KxShift) solver of the FEL runs in FP64
without space charge: modes -Nko..Nko, the carrier e^{-i kx0 x} taken out of
the deposits and put back in the gather, the transforms on kx_env = kx - kx0,
the operators and the complex PSATD tables on the full kx, the anti-echo
deposition filter (solvers.py:76-83, :164-176, :236-267, :660-759).

Extra key 'Precision': 'fp64' (default, the reference's arithmetic) or
'mixed': every field and FB-space array complex64, the Hankel, D+- and PSATD
tables built in FP64 and rounded to float32 (kx stays FP64), transforms and
operators as split-BF16 MFMA products with FP32 accumulation (six products of
three-piece BF16 splits: FP32 accuracy), or FP32 MFMA when the library is set
to that form.
"""
from __future__ import annotations

import numpy as np
from scipy.special import j1

from . import fimera as F
from . import operators as ops
from ._lib import lib
from .device import DeviceArray, call


class Solver:
    def __init__(self, solver_in):
        self.Args = solver_in
        self.frame = None     # device window frame (ChimeraRun.enable_device_frame), else None
        A = self.Args
        if "Features" not in A:
            A["Features"] = ()
        if "TimeActive" not in A:
            A["TimeInterval"] = (0.0, np.inf)        # solvers.py:48-49 (key kept as the reference sets it)
        A.setdefault("Precision", "fp64")
        if A["Precision"] not in ("fp64", "mixed"):
            raise ValueError(f"Precision {A['Precision']!r}: 'fp64' or 'mixed'")
        self.mixed = A["Precision"] == "mixed"
        self.env = "KxShift" in A
        if self.env:
            if self.mixed:
                raise NotImplementedError("the envelope (KxShift) solver runs in FP64 only")
            for feat in ("SpaceCharge", "StaticKick", "StillAsBackground"):
                # the reference builds no consistent envelope tables for these
                # (solvers.py:244-249 stores the complex coefficients into real ones)
                if feat in A["Features"]:
                    raise NotImplementedError(f"the envelope (KxShift) solver with {feat}")
        self.rt = np.float32 if self.mixed else np.float64          # real tables
        self.ct = np.complex64 if self.mixed else np.complex128     # fields

        A["Nko"] = A["MaxAzimuthMode"]
        A["leftX"], A["rightX"], A["lengthR"], A["dx"], A["dr"] = A["Grid"]
        A["dt"] = A["TimeStep"]
        A["dx_inv"], A["dr_inv"], A["dt_inv"] = 1.0 / A["dx"], 1.0 / A["dr"], 1.0 / A["dt"]

        nthrds = None
        if "Xchunked" in A:
            A["nthrds"] = A["Xchunked"][0]
            nthrds = A["nthrds"]
        A["Nx"], A["Xgrid"] = ops.x_grid(A["leftX"], A["rightX"], A["dx"], nthrds)
        if nthrds is not None:
            A["Nxchunk"] = A["Nx"] // nthrds
        if lib().chimera_fft_length_check(int(A["Nx"])) != 0:
            raise ValueError(lib().chimera_last_error().decode())
        A["leftX"] = A["Xgrid"][0]
        A["kx0"] = 2 * np.pi * A["KxShift"] if self.env else 0.0
        A["kx_env"] = 2 * np.pi * np.fft.fftfreq(A["Nx"], A["dx"])
        A["kx"] = A["kx0"] + A["kx_env"]
        A["dkx"] = (A["kx"][1] - A["kx"][0]) / (2.0 * np.pi)
        A["Nkr"], A["RgridFull"], A["lengthR"] = ops.r_grid(A["lengthR"], A["dr"])
        A["Nr"] = A["Nkr"] + 1

        so = ops.spectral_operators_env if self.env else ops.spectral_operators
        ext1 = so(A["Nx"], A["Nkr"], A["Nko"], A["lengthR"], A["kx"], A["RgridFull"], ext=1)
        ext0 = so(A["Nx"], A["Nkr"], A["Nko"], A["lengthR"], A["kx"], A["RgridFull"], ext=0)
        A["DpS2S"], A["DmS2S"] = ext1["DpS2S"], ext1["DmS2S"]
        A["InFull"], A["OutFull"] = ext0["In"], ext0["Out"]
        A["w"], A["kr"], A["kr_g"], A["kx_g"] = ext0["w"], ext0["kr"], ext0["kr_g"], ext0["kx_g"]
        A["idxM"] = ext0["idxM"] if self.env else np.arange(0, A["Nko"] + 1)
        A["Mtot"] = A["InFull"].shape[-1]
        A["PoissFact"] = 1.0 / A["w"] ** 2
        VGrid = 2 * np.pi * A["dx"] * A["dr"] * A["RgridFull"]
        A["VGrid"] = (A["RgridFull"] > 0.0) / VGrid
        A["InCurr"] = A["InFull"] * A["VGrid"][None, 1:, None]

        if "Rcut" in A:
            indRcut = int((A["RgridFull"] < A["Rcut"]).sum())
            A["Rgrid"] = A["RgridFull"][:indRcut]
            A["In"] = A["InFull"][:, :indRcut - 1]
            A["InCurr"] = A["InCurr"][:, :indRcut - 1]
            A["Out"] = A["OutFull"][:indRcut - 1]
            A["Rcut"] = A["Rgrid"].max()
        else:
            A["In"], A["Out"], A["Rgrid"] = A["InFull"], A["OutFull"], A["RgridFull"]
        A["Nr"] = A["Rgrid"].shape[0]
        A["lowerR"] = (A["Rgrid"] * (A["Rgrid"] >= 0)).min()
        A["upperR"] = A["Rgrid"].max()

        for kw in ("DpS2S", "DmS2S", "InCurr", "In", "Out"):      # solvers.py:152-157
            A[kw] = np.asfortranarray(np.swapaxes(A[kw], 0, 1))
        for kw in list(A.keys()):
            if type(A[kw]) == np.ndarray:
                A[kw] = np.asfortranarray(A[kw])

        A["FBDiff"] = [A[kw] for kw in ("DpS2S", "DmS2S", "kx")]
        A["DepProj"] = [A[kw] for kw in ("Rgrid", "dx_inv", "dr_inv")]
        if self.env:                                               # solvers.py:164-167
            filt_x = ops.antialias(A["kx_env"], A["kx"], A["kx0"], A["Nx"], A["Features"])
            A["DepFact"] = ops.dep_fact_env(A["Nx"], A["dr"], A["kr_g"], filt_x)
            A["DepProj"] += [A["kx0"]]
            kx_base = A["kx_env"]
        else:
            A["DepFact"] = ops.dep_fact(A["Nx"], A["dx"], A["dr"], A["kx_g"], A["kr_g"])
            kx_base = A["kx"]
        A["FBCurrIn"] = (kx_base, A["InCurr"])
        A["FBIn"] = (kx_base, A["In"])
        A["FBout"] = (kx_base, A["Out"])

        # ---- device copies of the operators; kx_base is the wave number of
        # the x transforms (kx_env for the envelope), kx that of the operators
        self.dev = {kw: DeviceArray.from_numpy(A[kw], self.rt)
                    for kw in ("DpS2S", "DmS2S", "In", "InCurr", "Out", "DepFact", "PoissFact")}
        for kw in ("kx", "Rgrid"):
            self.dev[kw] = DeviceArray.from_numpy(A[kw], np.float64)
        self.dev["kx_base"] = DeviceArray.from_numpy(kx_base, np.float64) if self.env else self.dev["kx"]

        # ---- device state (solvers.py:183-212)
        Nx, Nr, Nkr, M = A["Nx"], A["Nr"], A["Nkr"], A["Mtot"]
        sp, fb = (Nx, Nr, M), (Nx, Nkr, M)
        D = self.Data = {}
        ct = self.ct
        D["EB"] = DeviceArray(sp + (6,), ct)
        D["J"] = DeviceArray(sp + (3,), ct)
        D["scl_spc"] = DeviceArray(sp, ct)
        D["EG_fb"] = DeviceArray(fb + (6,), ct)
        D["vec_fb"] = DeviceArray(fb + (3,), ct)
        D["scl_fb"] = DeviceArray(fb, ct)
        D["J_fb"] = DeviceArray(fb + (3,), ct)
        D["B_fb"] = DeviceArray(fb + (3,), ct)
        if "SpaceCharge" in A["Features"] or "StaticKick" in A["Features"]:   # solvers.py:201-207
            D["Rho"] = DeviceArray(sp, ct)
            D["Rho_fb"] = DeviceArray(fb, ct)
            D["gradRho_fb_nxt"] = DeviceArray(fb + (3,), ct)
            D["gradRho_fb_prv"] = DeviceArray(fb + (3,), ct)
        if "StillAsBackground" in A["Features"]:
            D["BckGrndRho"] = DeviceArray(sp, ct)

        if "CoPropagative" in A:
            self.PSATD_coeffs(A["CoPropagative"])
        else:
            self.PSATD_coeffs()

    def fn(self, name):
        """Library entry point of this solver's precision."""
        return name + "_c8" if self.mixed else name

    # ------------------------------------------------------------------ geometry helpers
    @property
    def n_fb(self):
        A = self.Args
        return A["Nx"], A["Nkr"], A["Mtot"]

    @property
    def nko_env(self):
        """Highest envelope mode: the mode axis holds -nko_env..nko_env."""
        return (self.Args["Mtot"] - 1) // 2

    def geom_args(self):
        """(leftX, Rgrid, dx_inv, dr_inv, Rgrid(0), Rgrid(nr), nx, nr, nkO) for the resident particle calls."""
        A = self.Args
        R = A["Rgrid"]
        return (float(A["leftX"]), self.dev["Rgrid"].ptr, float(A["dx_inv"]), float(A["dr_inv"]), float(R[0]),
                float(R[-1]), A["Nx"] - 1, A["Nr"] - 1, A["Mtot"] - 1)

    def geom_args_env(self):
        """(leftX, Rgrid, dx_inv, dr_inv, Rgrid(0), Rgrid(nr), kx0, nx, nr, nkO) for the resident envelope calls."""
        g = self.geom_args()
        return g[:6] + (float(self.Args["kx0"]), g[6], g[7], self.nko_env)

    def frame_ptr(self, stage):
        """Device address of the window edge after the stage-1 (0) or the
        stage-2 (1) move of the current step when the run keeps its frame on
        the device, else None (the kernels then use Args['leftX'])."""
        return None if self.frame is None else self.frame.offset_ptr(stage)

    # ------------------------------------------------------------------ PSATD (solvers.py:227-299)
    def PSATD_coeffs(self, beta=1.0):
        A = self.Args
        if self.env:
            E, G = ops.psatd_tables_env(A["w"], A["kx_g"], A["TimeStep"], beta)
            self.Data["PSATD_E"] = DeviceArray.from_numpy(E, np.complex128)
            self.Data["PSATD_G"] = DeviceArray.from_numpy(G, np.complex128)
            return
        E, G = ops.psatd_tables(A["w"], A["TimeStep"], "SpaceCharge" in A["Features"])
        self.Data["PSATD_E"] = DeviceArray.from_numpy(E, self.rt)
        self.Data["PSATD_G"] = DeviceArray.from_numpy(G, self.rt)

    def maxwell_solver(self):
        D = self.Data
        Nx, Nkr, M = self.n_fb
        if self.env:
            call("chimera_maxwell_push_wo_spchrg_env", D["EG_fb"].ptr, D["J_fb"].ptr, D["PSATD_E"].ptr, D["PSATD_G"].ptr,
                 Nx, Nkr, M)
        elif "SpaceCharge" in self.Args["Features"]:
            call(self.fn("chimera_maxwell_push_with_spchrg"), D["EG_fb"].ptr, D["J_fb"].ptr, D["gradRho_fb_prv"].ptr,
                 D["gradRho_fb_nxt"].ptr, D["PSATD_E"].ptr, D["PSATD_G"].ptr, Nx, Nkr, M)
        else:
            call(self.fn("chimera_maxwell_push_wo_spchrg"), D["EG_fb"].ptr, D["J_fb"].ptr, D["PSATD_E"].ptr,
                 D["PSATD_G"].ptr, Nx, Nkr, M)

    def poiss_corr(self, poiss_corr_num=3):
        """solvers.py:301-331; each iteration is one fused graddiv + correction."""
        A, D = self.Args, self.Data
        if "NoPoissonCorrection" in A["Features"]:
            return
        Nx, Nkr, M = self.n_fb
        for _ in range(poiss_corr_num):
            if "SpaceCharge" in A["Features"]:
                call(self.fn("chimera_fb_graddiv_poisson"), D["J_fb"].ptr, D["gradRho_fb_prv"].ptr, D["gradRho_fb_nxt"].ptr,
                     float(A["dt_inv"]), self.dev["PoissFact"].ptr, self.dev["DpS2S"].ptr, self.dev["DmS2S"].ptr,
                     self.dev["kx"].ptr, Nx, Nkr, M - 1)
            else:
                D["vec_fb"].copy_(D["J_fb"])
                self.FBGradDiv()
                call(self.fn("chimera_omp_mult_vec"), D["vec_fb"].ptr, self.dev["PoissFact"].ptr, Nx, Nkr, M)
                call(self.fn("chimera_omp_add_vec"), D["J_fb"].ptr, D["vec_fb"].ptr, Nx, Nkr, M)

    def maxwell_solver_stat(self, px0):
        """solvers.py:333-362 (complex tables built on the host for each px0)."""
        A, D = self.Args, self.Data
        if "SpaceCharge" not in A["Features"] and "StaticKick" not in A["Features"]:
            return
        C1, C2 = ops.static_tables(A["kx_g"], A["w"], px0)
        c1, c2 = DeviceArray.from_numpy(C1, self.ct), DeviceArray.from_numpy(C2, self.ct)
        Nx, Nkr, M = self.n_fb
        call(self.fn("chimera_maxwell_init_push"), D["EG_fb"].ptr, D["J_fb"].ptr, D["gradRho_fb_nxt"].ptr, c1.ptr, c2.ptr, Nx,
             Nkr, M)

    def poiss_corr_stat(self, px0):
        """solvers.py:364-387: J_fb += (graddiv J_fb + DT gradRho_fb_nxt) PoissFact,
        DT = -i beta0 kx (nkx values, FP64 in both precisions)."""
        A, D = self.Args, self.Data
        if "NoPoissonCorrection" in A["Features"]:
            return
        DT = DeviceArray.from_numpy(-1.0j * px0 / np.sqrt(1 + px0 ** 2) * A["kx"], np.complex128)
        D["vec_fb"].copy_(D["J_fb"])
        self.FBGradDiv()
        Nx, Nkr, M = self.n_fb
        call(self.fn("chimera_poiss_corr_stat"), D["J_fb"].ptr, D["vec_fb"].ptr, D["gradRho_fb_nxt"].ptr, DT.ptr,
             self.dev["PoissFact"].ptr, Nx, Nkr, M)

    def field_drift(self, px0):
        """solvers.py:389-405: EG_fb *= exp(-0.5 i dt beta0 kx)."""
        Nx, Nkr, M = self.n_fb
        call(self.fn("chimera_field_drift"), self.Data["EG_fb"].ptr, self.dev["kx"].ptr,
             float(px0 / np.sqrt(1 + px0 ** 2)), float(self.Args["TimeStep"]), Nx, Nkr, M)

    # ------------------------------------------------------------------ FB transforms (solvers.py:407-465)
    def fb_curr_in(self):
        """J (unfolded deposit) -> J_fb, with the ghost fold and DepFact fused."""
        A, D = self.Args, self.Data
        Nx, Nkr, M = self.n_fb
        call(self.fn("chimera_fb_forward"), D["J_fb"].ptr, D["J"].ptr, float(A["leftX"]), self.dev["kx_base"].ptr,
             self.dev["InCurr"].ptr, self.dev["DepFact"].ptr, Nx, A["Nr"] - 1, M, Nkr, 3, 1, self.frame_ptr(0))

    def fb_dens_in(self):
        A, D = self.Args, self.Data
        Nx, Nkr, M = self.n_fb
        call(self.fn("chimera_fb_forward"), D["Rho_fb"].ptr, D["Rho"].ptr, float(A["leftX"]), self.dev["kx_base"].ptr,
             self.dev["InCurr"].ptr, self.dev["DepFact"].ptr, Nx, A["Nr"] - 1, M, Nkr, 1, 1, self.frame_ptr(1))

    def fb_fld_out(self):
        """EG_fb, B_fb -> EB with eb_correction (eb_correction_env for the
        envelope: ghost row +row 1 for a single mode, -row 1 otherwise) fused."""
        A, D = self.Args, self.Data
        Nx, Nkr, M = self.n_fb
        ghost = (3 if A["Nko"] == 0 else 4) if self.env else 2
        call(self.fn("chimera_fb_inverse"), D["EB"].ptr, D["EG_fb"].ptr, D["B_fb"].ptr, float(A["leftX"]),
             self.dev["kx_base"].ptr, self.dev["Out"].ptr, Nx, A["Nr"] - 1, M, Nkr, ghost, self.frame_ptr(1))

    # ------------------------------------------------------------------ FB operators (solvers.py:467-553)
    _ENV_OPS = {1: "chimera_fb_grad_env", 2: "chimera_fb_div_env"}

    def _op(self, op, out, inp):
        Nx, Nkr, M = self.n_fb
        if self.env:
            Dp, Dm, kx = self.dev["DpS2S"].ptr, self.dev["DmS2S"].ptr, self.dev["kx"].ptr
            if op == 3:
                call("chimera_fb_graddiv_env", out.ptr, Dp, Dm, kx, Nx, Nkr, self.nko_env, Nkr)
            else:
                call(self._ENV_OPS[op], out.ptr, inp.ptr, Dp, Dm, kx, Nx, Nkr, self.nko_env, Nkr)
            return
        call(self.fn("chimera_fb_op"), op, out.ptr, inp.ptr, self.dev["DpS2S"].ptr, self.dev["DmS2S"].ptr, self.dev["kx"].ptr,
             Nx, Nkr, M - 1)

    def FBGrad(self):
        self._op(1, self.Data["vec_fb"], self.Data["scl_fb"])

    def FBDiv(self):
        self._op(2, self.Data["scl_fb"], self.Data["vec_fb"])

    def FBGradDiv(self):
        self._op(3, self.Data["vec_fb"], self.Data["vec_fb"])

    def FBGradDens(self):
        self._op(1, self.Data["gradRho_fb_nxt"], self.Data["Rho_fb"])

    def G2B_FBRot(self):
        """B = rot(G) * PoissFact (fused)."""
        D = self.Data
        Nx, Nkr, M = self.n_fb
        G = D["EG_fb"].offset_ptr(3 * Nx * Nkr * M)
        if self.env:                                               # solvers.py:536-553 (fb_rot_env)
            call("chimera_fb_rot_env", D["B_fb"].ptr, G, self.dev["DpS2S"].ptr, self.dev["DmS2S"].ptr,
                 self.dev["kx"].ptr, Nx, Nkr, self.nko_env, Nkr)
            call("chimera_omp_mult_vec", D["B_fb"].ptr, self.dev["PoissFact"].ptr, Nx, Nkr, M)
            return
        call(self.fn("chimera_fb_rot_scaled"), D["B_fb"].ptr, G, self.dev["DpS2S"].ptr, self.dev["DmS2S"].ptr,
             self.dev["kx"].ptr, self.dev["PoissFact"].ptr, Nx, Nkr, M - 1)

    # ------------------------------------------------------------------ laser and absorbing layer
    def add_gauss_beam(self, S):
        """solvers.py:555-610: the profile is built on the host, every
        transform and operator runs on the device.  The envelope branch writes
        the Gaussian pulse straight into FB space, on mode m = 0, around the
        carrier k0."""
        A, D = self.Args, self.Data
        k0 = 2 * np.pi * S["k0"]
        a0 = 2 * np.pi * S["a0"]
        X_focus = S["x0"] - S["x_foc"]
        kx_g, w = A["kx_g"], A["w"][:, :, :, None]
        Nx, Nkr, M = self.n_fb
        if self.env:
            nk = A["Nko"]                                          # mode m = 0 on the -Nko..Nko axis
            kr0 = A["kr_g"][:, :, nk]
            e_s0 = a0 * 0.5 * np.pi ** 0.5 * S["Lx"] * S["LR"] ** 2 * A["dkx"] / A["lengthR"] ** 2
            vec = np.zeros((Nx, Nkr, M, 3), dtype=complex, order="F")
            vec[:, :, nk, 2] = e_s0 / j1(A["lengthR"] * kr0) ** 2 * np.exp(-1j * kx_g * S["x0"]) \
                * np.exp(-0.25 * (kx_g - k0) ** 2 * S["Lx"] ** 2 - 0.25 * kr0 ** 2 * S["LR"] ** 2)
            self._add_beam(vec, -1.0j * w, w, kx_g, X_focus)
            return
        Xgrid, Rgrid = A["Xgrid"], A["Rgrid"]
        scl_spc = np.zeros((Nx, A["Nr"], M), dtype=complex, order="F")
        scl_spc[:, :, 0] = a0 * np.cos(k0 * (Xgrid[:, None] - S["x0"])) * np.exp(
            -(Xgrid[:, None] - S["x0"]) ** 2 / S["Lx"] ** 2 - Rgrid[None, :] ** 2 / S["LR"] ** 2) \
            * (abs(Rgrid[None, :]) < 3.5 * S["LR"]) * (abs(Xgrid[:, None] - S["x0"]) < 3.5 * S["Lx"])
        scl_spc[:, 0, 0] = 0.0
        scl_fb = F.fb_scl_in(np.zeros((Nx, Nkr, M), complex, order="F"), scl_spc, A["leftX"], *A["FBIn"])
        vec = np.zeros((Nx, Nkr, M, 3), dtype=complex, order="F")
        vec[:, :, :, 2] = scl_fb / A["Nx"]
        DT = -1.0j * w * np.sign(kx_g[:, :, None, None] + (kx_g[:, :, None, None] == 0))
        self._add_beam(vec, DT, w, kx_g, X_focus)

    def _add_beam(self, vec, DT, w, kx_g, X_focus):
        """Divergence cleaning, free propagation from the focus over X_focus and
        accumulation into EG_fb (solvers.py:594-610)."""
        D = self.Data
        EE = self.div_clean_host(vec)
        GG = DT * EE
        vf = np.cos(w * X_focus) * EE + np.sin(w * X_focus) / w * GG
        GG = -w * np.sin(w * X_focus) * EE + np.cos(w * X_focus) * GG
        EE = vf * np.exp(1.0j * kx_g[:, :, None, None] * X_focus)
        GG = GG * np.exp(1.0j * kx_g[:, :, None, None] * X_focus)
        EG = D["EG_fb"].numpy().astype(np.complex128)
        EG[..., :3] += EE
        EG[..., 3:] += GG
        D["EG_fb"].upload(EG)
        D["vec_fb"].zero_()
        D["scl_fb"].zero_()

    def div_clean_host(self, vec):
        """solvers.py:644-650 on a host array (laser set-up)."""
        A = self.Args
        gd = F.fb_graddiv_env if self.env else F.fb_graddiv
        v = gd(np.array(vec, order="F"), *A["FBDiff"])
        v = F.omp_mult_vec(v, A["PoissFact"])
        return F.omp_add_vec(np.array(vec, order="F"), v)

    def get_damp_profile(self, Lf):
        return ops.damp_profile(Lf)

    def damp_field(self, config="left", damp_b=False):
        """solvers.py:619-632: the absorbing layer on E and G (B when damp_b)."""
        if config != "left":
            raise NotImplementedError("only the 'left' absorbing layer is defined")
        if damp_b:
            raise NotImplementedError("damp_b (filter B, rebuild G) is outside the PIC-step path")
        A, D = self.Args, self.Data
        prof = A["damp_profile"]
        if "damp_profile_dev" not in self.dev or self.dev["damp_profile_dev"].shape != prof.shape:
            self.dev["damp_profile_dev"] = DeviceArray.from_numpy(np.asarray(prof, np.float64), self.rt)
        Nx, Nkr, M = self.n_fb
        call(self.fn("chimera_fb_filter"), D["EG_fb"].ptr, float(A["leftX"]), self.dev["kx"].ptr,
             self.dev["damp_profile_dev"].ptr, len(prof), Nx, Nkr, M, 6)
