"""CPU restatement of the reference's Python driver (Solver, Specie, ChimeraRun)
for the plain PIC path, calling the C oracle (cpu_ref.c) where the reference
calls its f2py module.

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker; the product never imports it.

Covered: FixedCell and RandCell species (Still or pushed, external-field devices),
SpaceCharge / StillAsBackground / NoPoissonCorrection / StaticKick solvers, Rcut, moving
frames with AbsorbLayer / AddPlasma / IonsOnTop / Staged, and the
'PreDepositBackground' run feature (background deposited before the initial
half step), and the envelope (KxShift) solver without space charge: modes
-Nko..Nko, the env deposits and gather, complex PSATD tables, the
anti-echo deposition filter and the spectral laser set-up.  Arrays and call
order follow solvers.py, species.py and chimera_main.py line by line (cited
per method).
"""
from __future__ import annotations

import contextlib
import inspect

import numpy as np
from scipy.special import j1, jn, jn_zeros

from .fimera_ref import Fimera

ch = None


def _ch():
    global ch
    if ch is None:
        ch = Fimera("port")
    return ch


@contextlib.contextmanager
def rounding_variant():
    """Runs built and stepped inside this block use the port compiled with FMA
    contraction: the same operations in the same order, each multiply-add
    rounded once instead of twice.  Its spread from the plain port is the
    oracle's own rounding floor (which a shuffled particle order does not
    show for quantities whose rounding is amplified by cancellation)."""
    global ch
    saved = ch
    ch = Fimera("port_fma")
    try:
        yield
    finally:
        ch = saved


class Solver:
    """solvers.py:26-225."""

    def __init__(self, solver_in):
        c = _ch()
        A = self.Args = solver_in
        A.setdefault("Features", ())
        A["Nko"] = A["MaxAzimuthMode"]
        A["leftX"], A["rightX"], A["lengthR"], A["dx"], A["dr"] = A["Grid"]
        A["dt"] = A["TimeStep"]
        A["dx_inv"], A["dr_inv"], A["dt_inv"] = 1.0 / A["dx"], 1.0 / A["dr"], 1.0 / A["dt"]
        if "Xchunked" in A:
            A["nthrds"] = A["Xchunked"][0]
            A["Nxchunk"] = 2 * int(np.round(0.5 / A["nthrds"] / A["dx"] * (A["rightX"] - A["leftX"])))
            A["Nx"] = A["Nxchunk"] * A["nthrds"]
        else:
            A["Nx"] = int(2 * np.round(0.5 / A["dx"] * (A["rightX"] - A["leftX"])))
        A["Xgrid"] = A["rightX"] - A["dx"] * np.arange(A["Nx"])[::-1]
        A["leftX"] = A["Xgrid"][0]
        self.env = "KxShift" in A                                         # solvers.py:76-83
        A["kx0"] = 2 * np.pi * A["KxShift"] if self.env else 0.0
        A["kx_env"] = 2 * np.pi * np.fft.fftfreq(A["Nx"], A["dx"])
        A["kx"] = A["kx0"] + A["kx_env"]
        A["dkx"] = (A["kx"][1] - A["kx"][0]) / (2.0 * np.pi)
        A["Nkr"] = int(np.round(A["lengthR"] / A["dr"]))
        A["RgridFull"] = A["dr"] * (np.arange(A["Nkr"] + 1) - 0.5)
        A["lengthR"] = A["RgridFull"][-1] + A["dr"]
        _, _, A["DpS2S"], A["DmS2S"], _, _, _, _, _ = self.get_spectral_operators(ext=1)
        In, Out, _, _, A["w"], A["kr"], A["kr_g"], A["kx_g"], A["idxM"] = self.get_spectral_operators(ext=0)
        A["Mtot"] = In.shape[-1]
        A["PoissFact"] = 1.0 / A["w"] ** 2
        VG = 2 * np.pi * A["dx"] * A["dr"] * A["RgridFull"]
        A["InCurr"] = In * ((A["RgridFull"] > 0.0) / VG)[None, 1:, None]
        if "Rcut" in A:
            n = int((A["RgridFull"] < A["Rcut"]).sum())
            A["Rgrid"], A["In"], A["InCurr"], A["Out"] = A["RgridFull"][:n], In[:, :n - 1], A["InCurr"][:, :n - 1], Out[:n - 1]
        else:
            A["Rgrid"], A["In"], A["Out"] = A["RgridFull"], In, Out
        A["Nr"] = A["Rgrid"].shape[0]
        A["upperR"] = A["Rgrid"].max()
        for kw in ("DpS2S", "DmS2S", "InCurr", "In", "Out"):
            A[kw] = np.asfortranarray(np.swapaxes(A[kw], 0, 1))
        A["FBDiff"] = [A["DpS2S"], A["DmS2S"], A["kx"]]
        A["DepProj"] = [A["Rgrid"], A["dx_inv"], A["dr_inv"]]
        if self.env:                                                      # solvers.py:164-176
            filt_x = self.antialias()
            A["DepProj"] += [A["kx0"]]
            kx_base = A["kx_env"]
        else:
            filt_x = np.cos(0.5 * A["dx"] * A["kx_g"])[:, :, None] ** 2
            kx_base = A["kx"]
        A["DepFact"] = np.asfortranarray((2 * np.pi) ** 2 / A["Nx"] * filt_x * np.cos(0.5 * A["dr"] * A["kr_g"]) ** 2)
        A["FBCurrIn"], A["FBIn"], A["FBout"] = (kx_base, A["InCurr"]), (kx_base, A["In"]), (kx_base, A["Out"])
        sp, fb = (A["Nx"], A["Nr"], A["Mtot"]), (A["Nx"], A["Nkr"], A["Mtot"])
        z = lambda s: np.zeros(s, dtype=complex, order="F")
        D = self.Data = {"EB": z(sp + (6,)), "J": z(sp + (3,)), "scl_spc": z(sp), "EG_fb": z(fb + (6,)),
                         "vec_fb": z(fb + (3,)), "scl_fb": z(fb), "J_fb": z(fb + (3,)), "B_fb": z(fb + (3,))}
        if "SpaceCharge" in A["Features"] or "StaticKick" in A["Features"]:   # solvers.py:201-207
            D["Rho"], D["Rho_fb"] = z(sp), z(fb)
            D["gradRho_fb_nxt"], D["gradRho_fb_prv"] = z(fb + (3,)), z(fb + (3,))
        if "StillAsBackground" in A["Features"]:
            D["BckGrndRho"] = z(sp)
        self.PSATD_coeffs()
        self.c = c

    def antialias(self):                                                  # solvers.py:660-715
        A = self.Args
        kx_env, kx, kx0, Nx = A["kx_env"], A["kx"], A["kx0"], A["Nx"]
        ca = 0.85
        x = np.abs(kx_env) / np.abs(kx_env.max())
        band = ((x < ca) + (x >= ca) * np.cos(np.pi / 2 * (x - ca) / (1 - ca)) ** 2)[:, None, None]
        anti = np.ones_like(band)
        feats = A["Features"]
        if "NoAntiEcho" not in feats:
            ae = feats["AntiEchoStrength"] if "AntiEchoStrength" in feats else 2
            n_echo = int(np.abs(kx).max() / np.abs(kx_env).max()) + 1     # np.int in the reference
            echoes = np.abs(kx_env).max() / kx0 * np.arange(n_echo) - 1.0
            band_lims = np.array([kx.min() / kx0, kx.max() / kx0]) - 1.0
            k = 0
            for ce in echoes:
                if band_lims[0] < ce < band_lims[1]:
                    a = ae[k] if isinstance(ae, (list, tuple)) else ae
                    k += 1
                    if a <= 0:
                        continue
                    with np.errstate(divide="ignore", invalid="ignore"):
                        f = 1 - np.exp(-(kx / kx0 - 1 - ce) ** 2 / (a * (ce + 1.0) / Nx) ** 2)
                    anti = anti * f[:, None, None]
        return anti * band

    def get_spectral_operators(self, ext=0):
        """solvers.py:717-759."""
        A = self.Args
        Nx, Nkr, Nko, lengthR, kx, Rg = A["Nx"], A["Nkr"], A["Nko"], A["lengthR"], A["kx"], A["RgridFull"]
        if self.env:
            return self._spectral_operators_env(ext)
        Mmax, Mtot = Nko + ext, Nko + 1 + ext
        kr = np.zeros((Nkr, Mtot + 1))
        kr_g, w = np.zeros((Nx, Nkr, Mtot)), np.zeros((Nx, Nkr, Mtot))
        Out, In = np.zeros((Nkr, Nkr, Mtot)), np.zeros((Nkr, Nkr, Mtot))
        Dp, Dm = np.zeros((Nkr, Nkr, Mtot)), np.zeros((Nkr, Nkr, Mtot))
        for jm in range(Mmax + 2):
            kr[:, jm] = jn_zeros(jm, Nkr) / lengthR
        for jm in range(Mmax + 1):
            kr_g[:, :, jm], kx_g = np.meshgrid(kr[:, jm], kx)
            w[:, :, jm] = np.sqrt(kx_g ** 2 + kr_g[:, :, jm] ** 2)
            Out[:, :, jm] = jn(jm, Rg[1:, None] * kr[:, jm][None, :])
            In[:, :, jm] = np.linalg.inv(Out[:, :, jm])
            Dp[:, :, jm] = In[:, :, jm].dot(0.5 * kr[:, jm + 1][None, :] * jn(jm, Rg[1:, None] * kr[:, jm + 1][None, :]))
            km = kr[:, abs(jm - 1)][None, :]
            Dm[:, :, jm] = In[:, :, jm].dot(0.5 * km * jn(jm, Rg[1:, None] * km))
        return In, Out, Dp, Dm, w, kr, kr_g, kx_g, np.arange(0, Mmax + 1)

    def _spectral_operators_env(self, ext):
        """The KxShift branch of solvers.py:717-759: modes -Nko-ext..Nko+ext,
        negative m stored at NumPy's wrapped index, reordered when Nko > 0."""
        A = self.Args
        Nx, Nkr, Nko, lengthR, kx, Rg = A["Nx"], A["Nkr"], A["Nko"], A["lengthR"], A["kx"], A["RgridFull"]
        m0, m1, mt = -Nko - ext, Nko + ext, 2 * Nko + 1 + 2 * ext
        kr = np.zeros((Nkr, mt + 2))
        kr_g, w = np.zeros((Nx, Nkr, mt)), np.zeros((Nx, Nkr, mt))
        Out, In = np.zeros((Nkr, Nkr, mt)), np.zeros((Nkr, Nkr, mt))
        Dp, Dm = np.zeros((Nkr, Nkr, mt)), np.zeros((Nkr, Nkr, mt))
        for jm in range(m0 - 1, m1 + 2):
            kr[:, jm] = jn_zeros(jm, Nkr) / lengthR
        kx_g = None
        for jm in range(m0, m1 + 1):
            kr_g[:, :, jm], kx_g = np.meshgrid(kr[:, jm], kx)
            w[:, :, jm] = np.sqrt(kx_g ** 2 + kr_g[:, :, jm] ** 2)
            Out[:, :, jm] = jn(jm, Rg[1:, None] * kr[:, jm][None, :])
            In[:, :, jm] = np.linalg.inv(Out[:, :, jm])
            kp = kr[:, jm + 1][None, :]
            Dp[:, :, jm] = In[:, :, jm].dot(0.5 * kp * jn(jm, Rg[1:, None] * kp))
            km = kr[:, abs(jm - 1)][None, :]
            Dm[:, :, jm] = In[:, :, jm].dot(0.5 * km * jn(jm, Rg[1:, None] * km))
        if Nko > 0:
            Out, In, Dp, Dm, w, kr_g, kr = (np.concatenate((a[..., m0:], a[..., :m1 + 1]), axis=-1)
                                            for a in (Out, In, Dp, Dm, w, kr_g, kr))
        return In, Out, Dp, Dm, w, kr, kr_g, kx_g, np.arange(m0, m1 + 1)

    def PSATD_coeffs(self):
        """solvers.py:227-279."""
        A, w, dt = self.Args, self.Args["w"], self.Args["TimeStep"]
        if self.env:
            if "SpaceCharge" in A["Features"]:
                raise NotImplementedError("envelope solver with space charge")
            k = A.get("CoPropagative", 1.0) * A["kx_g"][:, :, None]
            E, G = np.zeros(w.shape + (3,), complex, order="F"), np.zeros(w.shape + (3,), complex, order="F")
            E[..., 0], E[..., 1] = np.cos(dt * w), np.sin(dt * w) / w
            G[..., 0], G[..., 1] = -w * np.sin(dt * w), np.cos(dt * w)
            with np.errstate(divide="ignore", invalid="ignore"):           # kx = 0 on the grid: 0/0, as in the reference
                E[..., 2] = np.exp(-0.5j * k * dt) * 1j * k / (w ** 2 - k ** 2) * (
                    1.0 - np.exp(1j * k * dt) * (w / (1j * k) * np.sin(dt * w) + np.cos(dt * w)))
                G[..., 2] = np.exp(-0.5j * k * dt) * w ** 2 / (w ** 2 - k ** 2) * (
                    1.0 + np.exp(1j * k * dt) * (1j * k / w * np.sin(dt * w) - np.cos(dt * w)))
            self.Data["PSATD_E"], self.Data["PSATD_G"] = E, G
            return
        n = 5 if "SpaceCharge" in A["Features"] else 3
        E, G = np.zeros(w.shape + (n,), order="F"), np.zeros(w.shape + (n,), order="F")
        E[..., 0], E[..., 1] = np.cos(dt * w), np.sin(dt * w) / w
        G[..., 0], G[..., 1] = -w * np.sin(dt * w), np.cos(dt * w)
        E[..., 2], G[..., 2] = -np.sin(dt * w) / w, 1 - np.cos(dt * w)
        if n == 5:
            E[..., 3] = (dt * w * np.cos(dt * w) - np.sin(dt * w)) / w ** 3 / dt
            E[..., 4] = (np.sin(dt * w) - dt * w) / w ** 3 / dt
            G[..., 3] = (1 - np.cos(dt * w) - dt * w * np.sin(dt * w)) / w ** 2 / dt
            G[..., 4] = (np.cos(dt * w) - 1) / w ** 2 / dt
        self.Data["PSATD_E"], self.Data["PSATD_G"] = E, G

    def maxwell_solver(self):                                             # solvers.py:281-299
        D = self.Data
        if "SpaceCharge" in self.Args["Features"]:
            D["EG_fb"] = self.c.maxwell_push_with_spchrg(D["EG_fb"], D["J_fb"], D["gradRho_fb_prv"],
                                                         D["gradRho_fb_nxt"], D["PSATD_E"], D["PSATD_G"])
        else:
            D["EG_fb"] = self.c.maxwell_push_wo_spchrg(D["EG_fb"], D["J_fb"], D["PSATD_E"], D["PSATD_G"])

    def poiss_corr(self, n=3):                                            # solvers.py:301-331
        A, D = self.Args, self.Data
        if "NoPoissonCorrection" in A["Features"]:
            return
        for _ in range(n):
            D["vec_fb"][:] = D["J_fb"]
            self.FBGradDiv()
            if "SpaceCharge" in A["Features"]:
                D["J_fb"] = self.c.poiss_corr(D["J_fb"], D["vec_fb"], D["gradRho_fb_prv"], D["gradRho_fb_nxt"],
                                              A["dt_inv"], A["PoissFact"])
            else:
                D["vec_fb"] = self.c.omp_mult_vec(D["vec_fb"], A["PoissFact"])
                D["J_fb"] = self.c.omp_add_vec(D["J_fb"], D["vec_fb"])

    def maxwell_solver_stat(self, px0):                                   # solvers.py:333-362
        A, D = self.Args, self.Data
        if "SpaceCharge" not in A["Features"] and "StaticKick" not in A["Features"]:
            return
        beta0 = px0 / np.sqrt(1 + px0 ** 2)
        w = A["w"]
        k = beta0 * A["kx_g"][:, :, None]
        C1 = np.zeros(w.shape + (2,), dtype=complex, order="F")
        C2 = np.zeros(w.shape + (2,), dtype=complex, order="F")
        C1[..., 0], C1[..., 1] = 1.0j * k / (w ** 2 - k ** 2), -1.0 / (w ** 2 - k ** 2)
        C2[..., 0], C2[..., 1] = w ** 2 / (w ** 2 - k ** 2), 1.0j * k / (w ** 2 - k ** 2)
        D["EG_fb"] = self.c.maxwell_init_push(D["EG_fb"], D["J_fb"], D["gradRho_fb_nxt"], C1, C2)

    def poiss_corr_stat(self, px0):                                       # solvers.py:364-387
        A, D = self.Args, self.Data
        if "NoPoissonCorrection" in A["Features"]:
            return
        DT = -1.0j * px0 / np.sqrt(1 + px0 ** 2) * A["kx"]
        D["vec_fb"][:] = D["J_fb"]
        self.FBGradDiv()
        D["J_fb"] = self.c.poiss_corr_stat(D["J_fb"], D["vec_fb"], D["gradRho_fb_nxt"], DT, A["PoissFact"])

    def field_drift(self, px0):                                           # solvers.py:389-405
        D = self.Data
        D["EG_fb"] = self.c.field_drift(D["EG_fb"], self.Args["kx"], px0 / np.sqrt(1 + px0 ** 2), self.Args["TimeStep"])

    def fb_curr_in(self):                                                 # solvers.py:407-420
        A, D = self.Args, self.Data
        D["J_fb"] = self.c.fb_vec_in(D["J_fb"], D["J"], A["leftX"], *A["FBCurrIn"])
        D["J_fb"] = self.c.omp_mult_vec(D["J_fb"], A["DepFact"])

    def fb_dens_in(self):                                                 # solvers.py:422-435
        A, D = self.Args, self.Data
        D["Rho_fb"] = self.c.fb_scl_in(D["Rho_fb"], D["Rho"], A["leftX"], *A["FBCurrIn"])
        D["Rho_fb"] = self.c.omp_mult_scl(D["Rho_fb"], A["DepFact"])

    def fb_fld_out(self):                                                 # solvers.py:450-465
        A, D = self.Args, self.Data
        D["EB"] = self.c.fb_eb_out(D["EB"], D["EG_fb"], D["B_fb"], A["leftX"], *A["FBout"])
        D["EB"] = self.c.eb_correction_env(D["EB"]) if self.env else self.c.eb_correction(D["EB"])

    def FBGradDiv(self):
        fn = self.c.fb_graddiv_env if self.env else self.c.fb_graddiv
        self.Data["vec_fb"] = fn(self.Data["vec_fb"], *self.Args["FBDiff"])

    def FBGradDens(self):
        fn = self.c.fb_grad_env if self.env else self.c.fb_grad
        self.Data["gradRho_fb_nxt"] = fn(self.Data["gradRho_fb_nxt"], self.Data["Rho_fb"], *self.Args["FBDiff"])

    def G2B_FBRot(self):                                                  # solvers.py:536-553
        D = self.Data
        fn = self.c.fb_rot_env if self.env else self.c.fb_rot
        D["B_fb"] = fn(D["B_fb"], np.asfortranarray(D["EG_fb"][:, :, :, 3:]), *self.Args["FBDiff"])
        D["B_fb"] = self.c.omp_mult_vec(D["B_fb"], self.Args["PoissFact"])

    def add_gauss_beam(self, S):                                          # solvers.py:555-610
        A, D = self.Args, self.Data
        k0, a0, X_focus = 2 * np.pi * S["k0"], 2 * np.pi * S["a0"], S["x0"] - S["x_foc"]
        kx_g, w = A["kx_g"], A["w"][:, :, :, None]
        if self.env:
            nk, kr0 = A["Nko"], A["kr_g"][:, :, A["Nko"]]
            e_s0 = a0 * 0.5 * np.pi ** 0.5 * S["Lx"] * S["LR"] ** 2 * A["dkx"] / A["lengthR"] ** 2
            D["vec_fb"][:, :, nk, 2] = e_s0 / j1(A["lengthR"] * kr0) ** 2 * np.exp(-1j * kx_g * S["x0"]) \
                * np.exp(-0.25 * (kx_g - k0) ** 2 * S["Lx"] ** 2 - 0.25 * kr0 ** 2 * S["LR"] ** 2)
            self._propagate_beam(-1.0j * w, w, kx_g, X_focus)
            return
        Xg, Rg = A["Xgrid"], A["Rgrid"]
        D["scl_spc"][:, :, 0] = a0 * np.cos(k0 * (Xg[:, None] - S["x0"])) * np.exp(
            -(Xg[:, None] - S["x0"]) ** 2 / S["Lx"] ** 2 - Rg[None, :] ** 2 / S["LR"] ** 2) \
            * (abs(Rg[None, :]) < 3.5 * S["LR"]) * (abs(Xg[:, None] - S["x0"]) < 3.5 * S["Lx"])
        D["scl_spc"][:, 0, 0] = 0.0
        D["scl_fb"] = self.c.fb_scl_in(D["scl_fb"], D["scl_spc"], A["leftX"], *A["FBIn"])
        D["vec_fb"][:, :, :, 2] = D["scl_fb"] / A["Nx"]
        DT = -1.0j * w * np.sign(kx_g[:, :, None, None] + (kx_g[:, :, None, None] == 0))
        self._propagate_beam(DT, w, kx_g, X_focus)

    def _propagate_beam(self, DT, w, kx_g, X_focus):                     # solvers.py:594-610
        D = self.Data
        EE = self.div_clean(D["vec_fb"].copy(order="F"))
        GG = DT * EE
        D["vec_fb"][:] = np.cos(w * X_focus) * EE + np.sin(w * X_focus) / w * GG
        GG = -w * np.sin(w * X_focus) * EE + np.cos(w * X_focus) * GG
        EE = D["vec_fb"].copy(order="F")
        EE *= np.exp(1.0j * kx_g[:, :, None, None] * X_focus)
        GG *= np.exp(1.0j * kx_g[:, :, None, None] * X_focus)
        D["EG_fb"][:, :, :, :3] += EE
        D["EG_fb"][:, :, :, 3:] += GG
        D["vec_fb"][:] = 0.0
        D["scl_fb"][:] = 0.0

    def div_clean(self, vec):                                             # solvers.py:644-650
        D = self.Data
        D["vec_fb"][:] = vec
        self.FBGradDiv()
        D["vec_fb"] = self.c.omp_mult_vec(D["vec_fb"], self.Args["PoissFact"])
        return self.c.omp_add_vec(vec, D["vec_fb"])

    def get_damp_profile(self, Lf):                                       # solvers.py:612-617
        N = int(Lf)
        i = np.arange(N)
        return (i >= 0.75 * N) * (0.5 - 0.5 * np.cos(np.pi * (i - 0.75 * N) / (0.25 * N))) ** 2

    def damp_field(self):                                                 # solvers.py:619-632 ('left', damp_b False)
        A, D = self.Args, self.Data
        for sl in (slice(0, 3), slice(3, 6)):
            D["EG_fb"][:, :, :, sl] = self.c.fb_filtr(np.asfortranarray(D["EG_fb"][:, :, :, sl]), A["leftX"], A["kx"],
                                                      A["damp_profile"], 0)


class Specie:
    """species.py:24-398 (FixedCell, RandCell)."""

    def __init__(self, specie_in):
        A = self.Args = specie_in
        for k, v in (("Features", ()), ("Charge", -1.0), ("Mass", 1.0), ("Density", 0.0),
                     ("MomentaMeans", (0.0, 0.0, 0.0)), ("MomentaSpreads", (0.0, 0.0, 0.0))):
            A.setdefault(k, v)
        self.Devices = A.get("Devices", ())
        A["leftX"], A["rightX"], A["lengthR"], A["dx"], A["dr"] = A["Grid"]
        A["push_fact"] = 2 * np.pi * A["Charge"] / A["Mass"]
        if "Xchunked" in A:
            A["nthrds"] = int(A["Xchunked"][0])
            A["Nx"] = 2 * A["nthrds"] * int(np.round(0.5 / A["dx"] / A["nthrds"] * (A["rightX"] - A["leftX"])))
        else:
            A["Nx"] = 2 * int(np.round(0.5 / A["dx"] * (A["rightX"] - A["leftX"])))
        A["Xgrid"] = A["rightX"] - A["dx"] * np.arange(A["Nx"])[::-1]
        A["leftX"] = A["Xgrid"][0]
        A["Nr"] = int(np.round(A["lengthR"] / A["dr"]))
        A["Rgrid"] = A["dr"] * (np.arange(A["Nr"]) - 0.5)
        A["lowerR"], A["upperR"] = (A["Rgrid"] * (A["Rgrid"] >= 0)).min(), A["Rgrid"].max()
        if "FixedCell" in A:                                              # species.py:95-111
            fc = A["FixedCell"]
            A["Num_p"] = int(np.prod(fc))
            pX, pR, pO = np.mgrid[1:fc[0]:fc[0] * 1j, 1:fc[1]:fc[1] * 1j, 1:fc[2]:fc[2] * 1j]
            A["Pax"] = (np.asfortranarray((pX.ravel() - 0.5) / fc[0]), np.asfortranarray((pR.ravel() - 0.5) / fc[1]),
                        np.asfortranarray(np.exp(2.0j * np.pi * (pO.ravel() - 1) / fc[2])))
        else:
            A["Num_p"] = int(A.get("RandCell", 0))
        A["wght0"] = A["Charge"] * A["Density"] * A["dr"] * A["dx"] * 2 * np.pi / A["Num_p"] if A["Num_p"] else 0.0
        self.Data = {"EB": np.zeros((6, 0), order="F"), "coords": np.zeros((3, 0), order="F"),
                     "weights": np.zeros((0,), order="F"), "coords_halfstep": np.zeros((3, 0), order="F"),
                     "momenta": np.zeros((3, 0), order="F")}

    def gen_parts(self, Domain=None, Xsteps=None, ProfileFunc=None):      # species.py:132-215
        A = self.Args
        Xgrid, Rgrid = A["Xgrid"], A["Rgrid"]
        if Domain is not None:
            pl, pr, plo, pup = Domain
            if A["leftX"] > pr or A["rightX"] < pl or pup < A["lowerR"] or plo > A["upperR"]:
                return [np.zeros((3, 0), order="F"), np.zeros((3, 0), order="F"), np.zeros(0)]
            Xgrid = Xgrid[(Xgrid < pl).sum() - 1:(Xgrid < pr).sum() + 1]
            irb = 0 if plo <= Rgrid.min() else (Rgrid < plo).sum() - 1
            ire = Rgrid.shape[0] if pup >= Rgrid.max() else (Rgrid < pup).sum() + 1
            Rgrid = Rgrid[irb:ire]
        elif Xsteps is not None:
            Xgrid = Xgrid[-2 * Xsteps:-Xsteps]
        coords = np.zeros((4, Xgrid.shape[0] * Rgrid.shape[0] * A["Num_p"]), order="F")
        if "FixedCell" in A:
            RandPackO = np.random.rand(Xgrid.shape[0], Rgrid.shape[0]).astype("d", order="F")
            coords, n = _ch().genparts(coords, Xgrid, Rgrid, RandPackO, *A["Pax"])
        else:
            coords, n = self.gen_randcell(coords, Xgrid, Rgrid)
        coords = coords[:, :n]
        if ProfileFunc is None:
            coords[-1] *= A["wght0"]
        else:
            na = len(inspect.signature(ProfileFunc).parameters)
            arg = {1: (coords[0],), 2: (coords[0], np.sqrt((coords[1:3] ** 2).sum(0))), 3: tuple(coords[0:3])}[na]
            coords[-1] *= A["wght0"] * ProfileFunc(*arg)
        rm = 2 * np.random.rand(3, n) - 1.0 if "FlatSpectrum" in A["Features"] else np.random.randn(3, n)
        mm, ms = A["MomentaMeans"], A["MomentaSpreads"]
        momenta = np.vstack([mm[c] + ms[c] * rm[c] for c in range(3)]).astype("d", order="F")
        return [np.asfortranarray(coords[0:3]), momenta, coords[-1].astype("d", order="F")]

    def gen_randcell(self, coords, Xgrid, Rgrid):                         # species.py:429-471
        """Num_p particles per cell (one cell per pair of neighbouring grid
        nodes), each axis stratified: the k-th particle sits in the k-th
        sub-interval, jittered by up to half of one, and the three axes are
        paired by three independent shuffles.  Draws from NumPy's global
        stream in the order the reference does: rand(3, Num_p), then the x, r
        and angle shuffles, cell by cell with r running fastest."""
        npc, dx, dr = self.Args["Num_p"], self.Args["dx"], self.Args["dr"]
        k = np.arange(npc)
        ip = 0
        for ix in range(Xgrid.shape[0] - 1):
            for ir in range(Rgrid.shape[0] - 1):
                jit = 2 * (np.random.rand(3, npc) - 0.5)
                xc = Xgrid[ix] + dx * (k + 0.5 * jit[0]) / npc
                rc = Rgrid[ir] + dr * (k + 0.5 * jit[1]) / npc
                oc = 2 * np.pi * (k + 0.5 * jit[2]) / npc
                np.random.shuffle(xc)
                np.random.shuffle(rc)
                np.random.shuffle(oc)
                coords[0, ip:ip + npc] = xc
                coords[1, ip:ip + npc] = rc * np.cos(oc)
                coords[2, ip:ip + npc] = rc * np.sin(oc)
                coords[3, ip:ip + npc] = rc
                ip += npc
        return coords, ip

    def beam_focus(self, x_foc):                                          # species.py:473-486
        D = self.Data
        gg = (1 + D["momenta"] ** 2).sum(0) ** 0.5
        pzmean = (D["momenta"][0] * D["weights"]).sum() / D["weights"].sum()
        D["coords"][0] = D["coords"][0] - x_foc * (D["momenta"][0] - pzmean) / pzmean / gg ** 2
        D["coords"][1:3] = D["coords"][1:3] - D["momenta"][1:3] / pzmean * x_foc
        D["coords_halfstep"][:] = D["coords"]

    def denoise(self, WaveNums2Kill):                                     # species.py:322-349
        D = self.Data
        for k_supp in WaveNums2Kill:
            for key in ("coords", "coords_halfstep"):
                mirror = D[key].copy(order="F")
                mirror[0] = mirror[0] + 0.5 / k_supp
                D[key] = np.asfortranarray(np.concatenate((D[key], mirror), axis=1))
            D["momenta"] = np.asfortranarray(np.concatenate((D["momenta"], D["momenta"]), axis=1))
            D["weights"] = np.concatenate((D["weights"], D["weights"]), axis=0)
            D["weights"] *= 0.5

    def add_particles(self, coords, momenta, weights):                    # species.py:217-244
        D = self.Data
        D["coords"] = np.asfortranarray(np.concatenate((D["coords"], coords), axis=1))
        D["coords_halfstep"] = np.asfortranarray(np.concatenate((D["coords_halfstep"], coords), axis=1))
        D["momenta"] = np.asfortranarray(np.concatenate((D["momenta"], momenta), axis=1))
        D["weights"] = np.concatenate((D["weights"], weights))

    def make_field(self):
        if "Still" in self.Args["Features"]:
            return
        self.Data["EB"] = np.zeros((6, self.Data["coords"].shape[1]), order="F")

    def make_device(self, i_step=0):
        if "Still" in self.Args["Features"]:
            return
        for dev in self.Devices:                                            # species.py:273-277
            fn = dev[0]
            name = (fn if isinstance(fn, str) else getattr(fn, "__name__", "")).lower()
            args = [np.asarray(a, dtype=float) if np.ndim(a) else float(a) for a in dev[1:]]
            self.Data["EB"] = getattr(_ch(), name)(self.Data["coords"], self.Data["EB"],
                                                   i_step * self.Args["TimeStep"], *args)

    def push_velocs(self, dt_frac=1.0):
        if self.Data["coords"].shape[-1] == 0 or "Still" in self.Args["Features"]:
            return
        self.Data["momenta"] = _ch().push_velocs(self.Data["momenta"], self.Data["EB"],
                                                 self.Args["push_fact"] * dt_frac * self.Args["TimeStep"])

    def push_coords(self, dt_frac=1.0):
        if self.Data["coords"].shape[1] == 0 or "Still" in self.Args["Features"]:
            return
        self.Data["coords"], self.Data["coords_halfstep"] = _ch().push_coords(
            self.Data["coords"], self.Data["momenta"], self.Data["coords_halfstep"], dt_frac * self.Args["TimeStep"])

    def chunk_and_damp(self, wind=None, SimDom=None, position="stag"):     # species.py:351-398 (stable order)
        D, A = self.Data, self.Args
        if D["coords"].shape[-1] == 0:
            return
        if wind is None:
            wind = {"Features": (), "AbsorbLayer": 0.0}
        if SimDom is None:
            SimDom = np.asfortranarray([A["leftX"] + wind["AbsorbLayer"] * A["dx"], A["rightX"], 0.0, A["upperR"] ** 2])
        key = "coords_halfstep" if ("Xchunked" in A and "NoSorting" not in wind["Features"] and position == "cntr") \
            else "coords"
        idx, n = _ch().sortpartsout(D[key], SimDom)
        idx = idx[:n].astype(np.int64)
        for comp in ("coords", "coords_halfstep", "momenta"):
            D[comp] = np.asfortranarray(D[comp][:, idx])
        D["weights"] = D["weights"][idx].copy()


class ChimeraRun:
    """chimera_main.py:23-318."""

    def shuffle_particles(self, seed):
        """Put every species' particles in another (seeded) order.  Not in the
        reference: the same run with its particles shuffled gives the oracle's
        own noise floor, the spread that the order of the floating-point sums
        over particles alone makes in fields and moments."""
        rng = np.random.RandomState(seed)
        for sp in self.Particles:
            n = sp.Data["coords"].shape[1]
            perm = rng.permutation(n)
            for key in ("coords", "coords_halfstep", "momenta", "weights", "EB"):
                v = sp.Data.get(key)
                if isinstance(v, np.ndarray) and v.shape[-1] == n:
                    sp.Data[key] = np.asfortranarray(v[..., perm])

    def __init__(self, SimComps):
        self.Particles = SimComps.get("Particles", ())
        self.Solvers = SimComps.get("Solvers", ())
        self.MovingFrames = SimComps["MovingFrames"] if "MovingFrames" in SimComps else ({},)
        self.Features = SimComps.get("Features", ())
        for wind in self.MovingFrames:
            for k, v in (("Features", ()), ("TimeActive", (0.0, np.inf)), ("Velocity", 1.0), ("TimeStep", 1.0),
                         ("Steps", 1)):
                wind.setdefault(k, v)
            f = 0.5 if "Staged" in wind["Features"] else 1.0
            wind["shiftX"] = f * wind["Velocity"] * wind["TimeStep"] * wind["Steps"]
        for solver in self.Solvers:
            for wind in self.MovingFrames:
                if "AbsorbLayer" in wind:
                    solver.Args["damp_profile"] = solver.get_damp_profile(wind["AbsorbLayer"])
        if "PreDepositBackground" in self.Features:
            for solver in self.Solvers:
                if "StillAsBackground" in solver.Args["Features"]:
                    self.dep_bg(solver)
        self.make_halfstep()

    def make_halfstep(self):                                              # chimera_main.py:61-80
        for sp in self.Particles:
            a = self.Solvers[0].Args
            SimDom = np.asfortranarray([a["leftX"], a["rightX"], 0, a["upperR"] ** 2])
            if SimDom[0] != SimDom[1]:
                sp.chunk_and_damp(SimDom=SimDom)
        self.project_current()
        self.project_density()
        for solver in self.Solvers:
            for sp in self.Particles:
                solver.maxwell_solver_stat(sp.Args["MomentaMeans"][0])
        self.project_fields()
        for sp in self.Particles:
            sp.make_device()
        for sp in self.Particles:
            sp.push_velocs(dt_frac=0.5)

    def make_step(self, istep):                                           # chimera_main.py:82-92
        self.frame_act(istep)
        for sp in self.Particles:
            sp.push_coords()
        self.chunk_particles(istep)
        self.project_current()
        self.frame_act(istep, "stage2")
        self.project_density()
        self.update_fields()
        self.project_fields()
        for sp in self.Particles:
            sp.make_device(istep)
        for sp in self.Particles:
            sp.push_velocs()

    def project_current(self):
        for solver in self.Solvers:
            D = solver.Data
            D["J"][:] = 0.0
            for sp in self.Particles:
                if sp.Data["coords"].shape[1] == 0 or "Still" in sp.Args["Features"]:
                    continue
                fn = _ch().dep_curr_env if solver.env else _ch().dep_curr    # chimera_main.py:153-181
                D["J"] = fn(sp.Data["coords_halfstep"], sp.Data["momenta"], sp.Data["weights"], D["J"],
                            solver.Args["leftX"], *solver.Args["DepProj"])
            solver.fb_curr_in()

    def update_fields(self):                                              # chimera_main.py:116-128
        for solver in self.Solvers:
            if "StaticKick" in solver.Args["Features"]:
                solver.Data["EG_fb"][:] = 0.0
                for sp in self.Particles:
                    PXmean = (sp.Data["momenta"][0] * sp.Data["weights"]).sum() / sp.Data["weights"].sum()
                    solver.poiss_corr_stat(PXmean)
                    solver.maxwell_solver_stat(PXmean)
                    solver.field_drift(PXmean)
            else:
                solver.poiss_corr()
                solver.maxwell_solver()

    def project_density(self):                                            # chimera_main.py:103-114
        for solver in self.Solvers:
            if "SpaceCharge" not in solver.Args["Features"] and "StaticKick" not in solver.Args["Features"]:
                continue
            if "StaticKick" not in solver.Args["Features"]:
                solver.Data["gradRho_fb_prv"][:] = solver.Data["gradRho_fb_nxt"]
            self.dep_dens(solver)
            solver.fb_dens_in()
            solver.FBGradDens()

    def dep_dens(self, solver):                                           # chimera_main.py:183-218
        D = solver.Data
        D["Rho"][:] = 0.0
        if "StillAsBackground" in solver.Args["Features"]:
            D["Rho"] += D["BckGrndRho"]
        component = "coords_halfstep" if "StaticKick" in solver.Args["Features"] else "coords"
        for sp in self.Particles:
            if sp.Data["coords"].shape[1] == 0 or "Still" in sp.Args["Features"]:
                continue
            fn = _ch().dep_dens_env if solver.env else _ch().dep_dens
            D["Rho"] = fn(sp.Data[component], sp.Data["weights"], D["Rho"], solver.Args["leftX"], *solver.Args["DepProj"])

    def dep_bg(self, solver):                                             # chimera_main.py:220-248
        D = solver.Data
        D["BckGrndRho"][:] = 0.0
        for sp in self.Particles:
            if sp.Data["coords"].shape[1] == 0 or "Still" not in sp.Args["Features"]:
                continue
            fn = _ch().dep_dens_env if solver.env else _ch().dep_dens
            D["BckGrndRho"] = fn(sp.Data["coords"], sp.Data["weights"], D["BckGrndRho"], solver.Args["leftX"],
                                 *solver.Args["DepProj"])

    def project_fields(self):                                             # chimera_main.py:130-151
        for sp in self.Particles:
            sp.make_field()
        for solver in self.Solvers:
            solver.G2B_FBRot()
            solver.fb_fld_out()
            for sp in self.Particles:
                if sp.Data["coords"].shape[1] == 0 or "Still" in sp.Args["Features"]:
                    continue
                fn = _ch().proj_fld_env if solver.env else _ch().proj_fld    # chimera_main.py:139-151
                sp.Data["EB"] = fn(sp.Data["coords"], sp.Data["weights"], solver.Data["EB"], sp.Data["EB"],
                                   solver.Args["leftX"], *solver.Args["DepProj"])

    def frame_act(self, istep, act="stage1"):                             # chimera_main.py:250-304
        for wind in self.MovingFrames:
            if istep < wind["TimeActive"][0] or istep > wind["TimeActive"][1]:
                continue
            if np.mod(istep - wind["TimeActive"][0], wind["Steps"]) != 0:
                continue
            if act == "stage1":
                if "AbsorbLayer" in wind and wind["AbsorbLayer"] > 0:
                    for solver in self.Solvers:
                        if "StaticKick" not in solver.Args["Features"]:
                            solver.damp_field()
                self.move_frame(wind)
                if "AddPlasma" in wind:
                    xs = int(wind["shiftX"] / wind["TimeStep"]) + 1
                    if "IonsOnTop" in wind["Features"]:
                        s1, s2 = self.Particles[:2]
                        parts = s1.gen_parts(Xsteps=xs, ProfileFunc=wind["AddPlasma"])
                        s1.add_particles(*parts)
                        parts[-1] = parts[-1] * -1
                        s2.add_particles(*parts)
                    else:
                        for sp in self.Particles:
                            sp.add_particles(*sp.gen_parts(Xsteps=xs, ProfileFunc=wind["AddPlasma"]))
                if "AbsorbLayer" in wind:
                    for sp in self.Particles:
                        sp.chunk_and_damp(wind=wind)
                for solver in self.Solvers:
                    if "StaticKick" in solver.Args["Features"]:
                        continue
                    if "SpaceCharge" in solver.Args["Features"]:
                        if "StillAsBackground" in solver.Args["Features"]:
                            self.dep_bg(solver)
                        self.dep_dens(solver)
            elif act == "stage2" and "Staged" in wind["Features"]:
                self.move_frame(wind)

    def move_frame(self, wind):
        for comp in tuple(self.Solvers) + tuple(self.Particles):
            comp.Args["Xgrid"] = comp.Args["Xgrid"] + wind["shiftX"]
            comp.Args["leftX"], comp.Args["rightX"] = comp.Args["Xgrid"][0], comp.Args["Xgrid"][-1]

    def chunk_particles(self, istep=0):                                   # chimera_main.py:306-318
        for sp in self.Particles:
            if "Xchunked" not in sp.Args or "NoSorting" in sp.Args["Features"]:
                continue
            if np.mod(istep, sp.Args["Xchunked"][1] + 1) != 0 or sp.Data["coords"].shape[-1] == 0:
                continue
            a = self.Solvers[0].Args
            sp.chunk_and_damp(SimDom=np.asfortranarray([a["leftX"], a["rightX"], 0, a["Rgrid"].max() ** 2]),
                              position="cntr")
