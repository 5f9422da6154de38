"""Pins of the oracle's host half (oracle/driver.py): the operator tables,
PSATD coefficients, laser initialisation, plasma generation and the beam
tools that the driver restates from solvers.py and species.py.  Each check
comes from the mathematics, not from the driver's own formulas, and is chosen
so that a plausible transcription mistake (a wrong mode index, a dropped
factor, a sign) fails it:

* Bessel roots: J_m(kr * lengthR) = 0 with lengthR = dr (Nkr + 1/2), the
  radial nodes at dr (i - 1/2) (solvers.py:87-91, :737);
* In Out = I per mode, and a single Bessel mode J_m(kr_j r) maps to the unit
  vector e_j (solvers.py:742-743);
* D+ and D- against the Bessel recurrences (d/dr + (m+1)/r) J_{m+1}(k r) =
  k J_m(k r) and (d/dr - (m-1)/r) J_{m-1}(k r) = -k J_m(k r), evaluated with
  scipy's derivative routine jvp, a different route from the driver's
  construction (solvers.py:744-747);
* the PSATD step (solvers.py:227-279 with maxwell_solvers.f90:18-60) against
  the exact solution of dE/dt = G - J, dG/dt = -w^2 E - grad rho(t) with J
  constant and grad rho linear in t over the step, from the matrix
  exponential of the augmented linear system;
* add_gauss_beam (solvers.py:555-610): the on-axis envelope follows the
  Gaussian-beam law a0 / sqrt(1 + (X / (pi w0^2))^2) at distance X from the
  focus (doc/fel-lpa-demo.ipynb:812), before and after free propagation
  toward the focus;
* gen_parts (species.py:132-215, particle_tools.f90:84-128): the lattice in
  closed form and the weights through the exact charge integral
  sum w = q n pi R^2 L of the filled cells;
* denoise (species.py:322-349): the Fourier component of the weights at
  every killed wavenumber vanishes; beam_focus (species.py:473-486): the
  transverse emittance is unchanged and a ballistic drift of x_foc restores
  the transverse positions.

The discrete operators do not satisfy div(rot v) = 0 or rot(grad phi) = 0
exactly (D+ and D- project between truncated Bessel bases of different
orders; the residual is O(1) for generic fields), so those identities are not
used.
"""
import numpy as np
import pytest
from scipy.linalg import expm
from scipy.special import jv, jvp

from oracle import driver as O


def solver(nko=2, grid=(-6.4, 0.0, 9.6, 0.1, 0.3), features=("SpaceCharge",), dt=0.1):
    return O.Solver({"Grid": grid, "TimeStep": dt, "MaxAzimuthMode": nko, "Features": features})


def test_bessel_roots_and_radial_grid():
    s = solver()
    A = s.Args
    In, Out, Dp, Dm, w, kr, kr_g, kx_g, idxM = s.get_spectral_operators(ext=1)
    Nkr, dr = A["Nkr"], A["dr"]
    assert A["lengthR"] == pytest.approx(dr * (Nkr + 0.5), rel=1e-15)
    assert np.allclose(A["RgridFull"], dr * (np.arange(Nkr + 1) - 0.5), rtol=0, atol=1e-15)
    for m in range(kr.shape[1]):
        assert np.abs(jv(m, kr[:, m] * A["lengthR"])).max() < 1e-13
        assert np.all(np.diff(kr[:, m]) > 0) and kr[0, m] > 0
        # consecutive roots of J_m are about pi apart (McMahon)
        assert np.allclose(np.diff(kr[-5:, m]) * A["lengthR"], np.pi, rtol=2e-3)
    assert np.allclose(w, np.sqrt(kx_g[:, :, None] ** 2 + kr_g ** 2), rtol=1e-15)


def test_hankel_pair_and_single_mode():
    s = solver(nko=3)
    In, Out, *_ = s.get_spectral_operators(ext=0)
    A = s.Args
    r = A["RgridFull"][1:]
    kr = s.get_spectral_operators(ext=0)[5]
    for m in range(In.shape[-1]):
        assert np.abs(In[:, :, m] @ Out[:, :, m] - np.eye(A["Nkr"])).max() < 1e-12
        j = 7
        e = In[:, :, m] @ jv(m, kr[j, m] * r)
        ref = np.zeros(A["Nkr"])
        ref[j] = 1.0
        assert np.abs(e - ref).max() < 1e-12


def test_dpm_against_bessel_recurrences():
    s = solver(nko=3)
    In, Out, Dp, Dm, w, kr, kr_g, kx_g, idxM = s.get_spectral_operators(ext=1)
    r = s.Args["RgridFull"][1:, None]
    rng = np.random.default_rng(1)
    n = s.Args["Nkr"]
    for m in range(In.shape[-1]):
        c = rng.standard_normal(n) * np.exp(-np.arange(n) / 5)
        k = kr[:, m + 1][None, :]
        g = (c * (k * jvp(m + 1, r * k) + (m + 1) / r * jv(m + 1, r * k))).sum(1)
        ref = In[:, :, m] @ g
        assert np.abs(2 * Dp[:, :, m] @ c - ref).max() <= 1e-12 * np.abs(ref).max(), ("D+", m)
        if m >= 1:
            k = kr[:, m - 1][None, :]
            g = (c * (k * jvp(m - 1, r * k) - (m - 1) / r * jv(m - 1, r * k))).sum(1)
            ref = In[:, :, m] @ g
            assert np.abs(-2 * Dm[:, :, m] @ c - ref).max() <= 1e-12 * np.abs(ref).max(), ("D-", m)
    # mode 0: D- uses the roots of J_1 (|m - 1| = 1) and stands for the m = -1
    # neighbour J_{-1} = -J_1: (d/dr + 1/r) J_1(k r) = k J_0(k r), so
    # 2 D-_0 c = In_0 [(d/dr + 1/r) sum_k c_k J_1(k r)]
    c = rng.standard_normal(n) * np.exp(-np.arange(n) / 5)
    k = kr[:, 1][None, :]
    g = (c * (k * jvp(1, r * k) + 1.0 / r * jv(1, r * k))).sum(1)
    ref = In[:, :, 0] @ g
    assert np.abs(2 * Dm[:, :, 0] @ c - ref).max() <= 1e-12 * np.abs(ref).max()


def test_psatd_step_is_exact():
    """One PSATD step of the driver's tables through the oracle kernel equals
    the exact evolution over dt of E' = G - J, G' = -w^2 E - g(t) with
    g(t) = g_n + (g_{n+1} - g_n) t / dt (the fields' sources held as the
    scheme assumes), for every (kx, kr, m) and component."""
    s = solver(nko=1, features=("SpaceCharge",), dt=0.1)
    A, D = s.Args, s.Data
    rng = np.random.default_rng(3)
    shp = D["EG_fb"].shape
    cplx = lambda sh: np.asfortranarray(rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    EG, J = cplx(shp), cplx(shp[:3] + (3,))
    gn, gp = cplx(shp[:3] + (3,)), cplx(shp[:3] + (3,))
    out = s.c.maxwell_push_with_spchrg(EG.copy(order="F"), J, gn, gp, D["PSATD_E"], D["PSATD_G"])
    dt = A["dt"]
    w = A["w"]
    idx = [(0, 0, 0), (5, 3, 1), (17, 10, 0), (31, 31, 1), (40, 2, 1)]
    for (ix, ik, m) in idx:
        ww = w[ix, ik, m]
        # state (E, G, J, g_n, g_{n+1}, s) with s(t) = (g_{n+1} - g_n) t / dt:
        # E' = G - J, G' = -w^2 E - g_n - s, s' = (g_{n+1} - g_n) / dt, the rest constant
        M = np.zeros((6, 6))
        M[0, 1], M[0, 2] = 1.0, -1.0
        M[1, 0], M[1, 3], M[1, 5] = -ww ** 2, -1.0, -1.0
        M[5, 4], M[5, 3] = 1.0 / dt, -1.0 / dt
        P = expm(M * dt)
        for c in range(3):
            y0 = np.array([EG[ix, ik, m, c], EG[ix, ik, m, c + 3], J[ix, ik, m, c], gn[ix, ik, m, c],
                           gp[ix, ik, m, c], 0.0])
            y1 = P @ y0
            ref = np.array([y1[0], y1[1]])
            got = np.array([out[ix, ik, m, c], out[ix, ik, m, c + 3]])
            scale = np.abs(y0).max() * max(1.0, ww) * max(1.0, 1.0 / ww)
            assert np.abs(got - ref).max() <= 1e-12 * scale, (ix, ik, m, c, got, ref)


def test_psatd_conserves_vacuum_energy():
    s = solver(nko=1, features=(), dt=0.1)
    A, D = s.Args, s.Data
    rng = np.random.default_rng(4)
    EG = np.asfortranarray(rng.standard_normal(D["EG_fb"].shape) + 1j * rng.standard_normal(D["EG_fb"].shape))
    J = np.zeros(D["J_fb"].shape, complex, order="F")
    w = A["w"][..., None]
    e0 = np.abs(EG[..., :3]) ** 2 + np.abs(EG[..., 3:]) ** 2 / w ** 2
    out = EG
    for _ in range(25):
        out = s.c.maxwell_push_wo_spchrg(out, J, D["PSATD_E"], D["PSATD_G"])
    e1 = np.abs(out[..., :3]) ** 2 + np.abs(out[..., 3:]) ** 2 / w ** 2
    assert np.abs(e1 - e0).max() <= 1e-12 * e0.max()


def _laser_envelope(delta, steps=0, w0=3.0):
    """The on-axis (first radial node) envelope peak of E_z of a laser focused
    delta behind its start (x_foc = x0 - delta), after `steps` vacuum PSATD
    steps of dt = 0.05, and the distance from the focus it has reached."""
    s = solver(nko=1, grid=(-30.0, 10.0, 18.0, 0.05, 0.3), features=(), dt=0.05)
    s.add_gauss_beam({"a0": 1.0, "k0": 1.0, "x0": -10.0, "x_foc": -10.0 - delta, "Lx": 3.0, "LR": w0})
    for _ in range(steps):
        s.maxwell_solver()
    s.G2B_FBRot()
    s.fb_fld_out()
    ez = s.Data["EB"][:, 1, 0, 2]
    F = np.fft.fft(ez)
    k = np.fft.fftfreq(ez.size)
    F[k < 0] = 0.0
    F[k > 0] *= 2.0
    env = np.abs(np.fft.ifft(F)).max()
    X = delta + steps * 0.05
    zr = np.pi * w0 ** 2
    wx = w0 * np.sqrt(1 + (X / zr) ** 2)
    r1 = s.Args["Rgrid"][1]
    # mode 0 of a real field carries it times 1/(2 pi) after eb_correction (grid_deps.f90:229-246),
    # and add_gauss_beam scales a0 by 2 pi (solvers.py:571): the envelope is a0 times the law
    law = 1.0 / np.sqrt(1 + (X / zr) ** 2) * np.exp(-r1 ** 2 / wx ** 2)
    return env, law


@pytest.mark.slow
def test_gauss_beam_amplitude_law():
    for delta in (0.0, np.pi * 9.0, -14.0):
        env, law = _laser_envelope(delta)
        assert env == pytest.approx(law, rel=1e-2), delta
    # launched 14 before the focus, 200 steps (10 wavelengths) of vacuum propagation toward it
    env, law = _laser_envelope(-14.0, steps=200)
    assert env == pytest.approx(law, rel=1e-2)
    env_far, law_far = _laser_envelope(-14.0)
    assert env > env_far * 1.05          # it converged


def test_gen_parts_lattice_and_charge():
    g = (-3.2, 0.0, 2.4, 0.1, 0.3)
    sp = O.Specie({"Grid": g, "TimeStep": 0.1, "Density": 0.004, "FixedCell": (2, 3, 4), "Charge": -1.0,
                   "MomentaSpreads": (0.0, 0.0, 0.0)})
    np.random.seed(8)
    coords, mom, w = sp.gen_parts()
    A = sp.Args
    nx, nr = A["Xgrid"].size, A["Rgrid"].size
    assert w.size == (nx - 1) * (nr - 1) * 24
    x, r = coords[0], np.hypot(coords[1], coords[2])
    ix = np.floor((x - A["Xgrid"][0]) / A["dx"] + 1e-9).astype(int)
    ir = np.floor(r / A["dr"] + 1e-9).astype(int)
    # lattice offsets inside each cell: (i - 1/2)/n of the cell in x and r
    fx = np.round((x - A["Xgrid"][0]) / A["dx"] - ix, 12)
    fr = np.round(r / A["dr"] - ir, 12)
    assert set(np.unique(fx)) == {0.25, 0.75}
    assert np.allclose(np.unique(fr), [1 / 6, 0.5, 5 / 6])
    # within a cell the azimuths are one random rotation plus quarter turns
    # (y = r sin, z = r cos): their residues modulo a quarter turn agree
    th = np.arctan2(coords[1], coords[2])
    cell = ix * 1000 + ir
    for c in np.unique(cell):
        t = np.mod(th[cell == c], np.pi / 2)
        assert t.size == 24
        d = np.abs(t - t[0])
        assert np.all(np.minimum(d, np.pi / 2 - d) < 1e-12), c
    # total charge = q n times the volume of the filled cells (exact for the midpoint lattice in r)
    L = (nx - 1) * A["dx"]
    R = (nr - 1) * A["dr"]
    assert w.sum() == pytest.approx(A["Charge"] * A["Density"] * np.pi * R ** 2 * L, rel=1e-13)


def test_denoise_kills_the_wavenumbers():
    sp = O.Specie({"Grid": (-3.2, 0.0, 2.4, 0.1, 0.3), "TimeStep": 0.1, "Density": 0.004, "FixedCell": (1, 1, 2),
                   "MomentaSpreads": (0.1, 0.1, 0.1)})
    np.random.seed(2)
    sp.add_particles(*sp.gen_parts())
    rng = np.random.default_rng(0)
    sp.Data["coords"][0] += rng.uniform(-0.05, 0.05, sp.Data["coords"].shape[1])   # shot noise
    sp.Data["coords_halfstep"][:] = sp.Data["coords"]
    n0, q0 = sp.Data["weights"].size, sp.Data["weights"].sum()
    ks = (1.0, 2.5)
    sp.denoise(ks)
    D = sp.Data
    assert D["weights"].size == 4 * n0 and D["weights"].sum() == pytest.approx(q0, rel=1e-14)
    for k in ks:
        bunching = (D["weights"] * np.exp(2j * np.pi * k * D["coords"][0])).sum()
        assert abs(bunching) <= 1e-12 * abs(q0)
    assert np.array_equal(D["coords"], D["coords_halfstep"])


def test_beam_focus_is_a_ballistic_drift():
    sp = O.Specie({"Grid": (-3.2, 0.0, 2.4, 0.1, 0.3), "TimeStep": 0.1, "Density": 0.004, "FixedCell": (2, 2, 4),
                   "MomentaMeans": (200.0, 0.0, 0.0), "MomentaSpreads": (0.2, 0.3, 0.3)})
    np.random.seed(5)
    sp.add_particles(*sp.gen_parts())
    D = sp.Data
    y0, u, w = D["coords"][1:3].copy(), D["momenta"].copy(), D["weights"]

    def emit(y, p):
        return np.sqrt((w * y ** 2).sum() * (w * p ** 2).sum() - (w * y * p).sum() ** 2) / w.sum()

    e0 = [emit(y0[c], u[1 + c]) for c in range(2)]
    x_foc = 40.0
    sp.beam_focus(x_foc)
    pz = (u[0] * w).sum() / w.sum()
    for c in range(2):
        assert emit(D["coords"][1 + c], u[1 + c]) == pytest.approx(e0[c], rel=1e-10)
        # drifting x_foc at the mean longitudinal momentum brings the particles back
        assert np.allclose(D["coords"][1 + c] + u[1 + c] / pz * x_foc, y0[c], rtol=0, atol=1e-12)
    assert np.array_equal(D["coords"], D["coords_halfstep"])
