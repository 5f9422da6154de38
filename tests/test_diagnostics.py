"""Beam envelopes (get_beam_envelops, diagnostics.py:177-207) from sums
reduced on the device.

CPU: the oracle's restatement is pinned by a two-particle closed form and by
numpy's weighted average and covariance; the host formula that turns the 13
sums into the envelope is checked against the oracle on random beams.
GPU: the device sums of a bunch and of an lpa_cfg0 run, in FP64 and mixed
precision, against the oracle on the same particles copied to the host
(1e-9 relative: the sums differ in order only, and the rms and emittance
formulas subtract), the same bits on a second call, zeros for no particles,
and a refusal of host particle arrays.
"""

import numpy as np
import pytest

from oracle import diagnostics_ref as D


def test_oracle_envelope_closed_form():
    x = np.array([[1.0, 3.0], [0.0, 0.0], [2.0, 2.0]])
    u = np.array([[2.0, 2.0], [1.0, -1.0], [0.0, 4.0]])
    w = np.array([1.0, 1.0])
    env = D.beam_envelope(x, u, w)
    assert np.allclose(env[0], [2.0, 0.0, 2.0])
    assert np.allclose(env[1], [1.0, 0.0, 0.0])
    # x: sqrt(<x^2><px^2> - <x px>^2) = sqrt(5*4 - 4^2); y: 0; z: sqrt(4*8 - 4^2)
    assert np.allclose(env[2], [2.0, 0.0, 4.0])


def test_oracle_envelope_against_numpy():
    r = np.random.default_rng(11)
    x, u = r.normal(size=(3, 5000)) * [[3.0], [0.5], [0.7]] + [[-4.0], [0.1], [0.0]], r.normal(size=(3, 5000))
    w = r.uniform(0.1, 2.0, 5000)
    env = D.beam_envelope(x, u, w)
    for c in range(3):
        assert np.isclose(env[0, c], np.average(x[c], weights=w), rtol=1e-12)
        assert np.isclose(env[1, c], np.sqrt(np.cov(x[c], aweights=w, bias=True)), rtol=1e-10)


def test_envelope_from_sums_matches_oracle():
    from chimera_amd.diagnostics import Diagnostics
    r = np.random.default_rng(12)
    x, u, w = r.normal(size=(3, 3000)) - 2.0, r.normal(size=(3, 3000)) + 5.0, r.uniform(0.5, 1.5, 3000)
    env = Diagnostics.envelope_from_sums(D.moment_sums(x, u, w))
    assert np.allclose(env, D.beam_envelope(x, u, w), rtol=1e-12, atol=0)


def _runs(precision):
    import chimera_amd as G
    from chimera_amd import workloads as W
    bunch = W.with_precision(W.beam(Nx=256, Nkr=48, gamma=150.0, sig_x=2.0, sig_r=1.0, emittance=0.5, n_target=5e4,
                                    K=2.0, lambda_u=150.0, steps=6), precision)
    runs = [W.build_beam(G, bunch, 3), W.build(G, W.with_precision(W.lpa_cfg0(), precision), 5)]
    for run in runs:
        for i in range(1, 5):
            run.make_step(i)
    return runs


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_envelopes_match_oracle(precision):
    from chimera_amd.diagnostics import Diagnostics
    for run in _runs(precision):
        dg = Diagnostics(run)
        envs = dg.get_beam_envelops()
        live = [sp for sp in run.Particles if "Still" not in sp.Args["Features"]]
        assert len(envs) == len(live) >= 1
        for env, sp in zip(envs, live):
            xh = sp.host("coords_halfstep").astype(np.float64)
            u, w = sp.host("momenta").astype(np.float64), sp.host("weights").astype(np.float64)
            ref = D.beam_envelope(xh, u, w)
            # means relative to |mean| + rms (a centred beam has means near 0)
            scale = np.abs(ref) + np.r_[[ref[1]], np.zeros((2, 3))]
            err = np.max(np.abs(env - ref) / np.maximum(scale, 1e-300))
            print(f"case: {precision} {sp.np} particles, envelope rel err {err:.2e}", flush=True)
            assert err < 1e-9, (env, ref)
            s1, s2 = dg.beam_moment_sums(sp), dg.beam_moment_sums(sp)
            assert np.array_equal(s1, s2)
            # each sum against the sum of its terms' magnitudes (order of summation only)
            mag = D.moment_sums(np.abs(xh), np.abs(u), np.abs(w))
            assert np.all(np.abs(s1 - D.moment_sums(xh, u, w)) <= 1e-12 * mag), (s1, D.moment_sums(xh, u, w))


@pytest.mark.gpu
def test_moments_edge_cases():
    from chimera_amd._lib import lib
    from chimera_amd.device import DeviceArray
    L = lib()
    out = np.full(13, 7.0)
    x, u, w = DeviceArray((8, 3), np.float64), DeviceArray((8, 3), np.float64), DeviceArray((8,), np.float64)
    assert L.chimera_beam_moments(x.ptr, u.ptr, w.ptr, 8, 0, out.ctypes.data) == 0
    assert not np.any(out)
    host = np.zeros(24)
    assert L.chimera_beam_moments(host.ctypes.data, u.ptr, w.ptr, 8, 8, out.ctypes.data) != 0
    assert L.chimera_beam_moments(x.ptr, u.ptr, w.ptr, 4, 8, out.ctypes.data) != 0


# ------------------------------------------------------------------ EM energy per kx (nrg_out)

def test_oracle_nrg_out_single_element():
    N, K, M = 8, 3, 2
    EF = np.random.default_rng(3).uniform(1.0, 2.0, (N, K, M))
    for x0 in range(N):
        EG = np.zeros((N, K, M, 6), complex)
        EG[x0, 1, 1, 2] = 3.0 - 4.0j           # |.|^2 = 25, an E component
        EG[x0, 2, 0, 4] = 100.0                 # a G component: no contribution
        dat = D.nrg_out(EG, EF)
        pos = x0 - (N // 2 + 1) if x0 >= N // 2 + 1 else x0 + N - (N // 2 + 1)
        expect = np.zeros(N)
        expect[pos] = 25.0 * EF[x0, 1, 1]
        assert np.allclose(dat, expect, rtol=1e-15, atol=0)


def test_oracle_energy_fact_bessel_normalisation():
    """EnergyFact / prefactor = J_{|m|+1}(j_{m,k})^2 = 2 * int_0^1 J_m(j_{m,k} rho)^2 rho d rho:
    the index |m|+1 and the argument kr*lengthR of solvers.py:145-147."""
    from scipy.integrate import quad
    from scipy.special import jn, jn_zeros
    R, dkx, nk = 3.0, 0.25, 4
    idxM = np.arange(3)
    kr_g = np.stack([jn_zeros(m, nk) / R for m in idxM], -1)[None, :, :]
    EF = D.energy_fact(idxM, kr_g, R, dkx)
    from scipy.constants import c, e, epsilon_0, m_e
    pref = 0.5 * 0.5e-2 * (m_e * c ** 2 / e) ** 2 * (4 * np.pi * epsilon_0) * R ** 2 / dkx
    for m in idxM:
        for k, z in enumerate(jn_zeros(m, nk)):
            integral = quad(lambda rho: jn(m, z * rho) ** 2 * rho, 0.0, 1.0, epsabs=1e-14, epsrel=1e-12)[0]
            assert np.isclose(EF[0, k, m] / pref, 2.0 * integral, rtol=1e-9)


def test_energy_fact_table_matches_oracle():
    from chimera_amd import operators as ops
    from chimera_amd.diagnostics import Diagnostics
    Nx, Nkr, Nko, R = 64, 16, 2, 4.8
    kx = 2 * np.pi * np.fft.fftfreq(Nx, 0.1)
    Nkr, Rg, R = ops.r_grid(R, 0.3)
    ext0 = ops.spectral_operators(Nx, Nkr, Nko, R, kx, Rg, ext=0)
    A = {"idxM": np.arange(Nko + 1), "lengthR": R, "dkx": (kx[1] - kx[0]) / (2 * np.pi), "kr_g": ext0["kr_g"]}
    assert np.array_equal(Diagnostics.energy_fact_table(A), D.energy_fact(A["idxM"], A["kr_g"], R, A["dkx"]))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_nrg_out_matches_oracle(precision, tmp_path):
    import chimera_amd as G
    from chimera_amd import workloads as W
    from chimera_amd.diagnostics import Diagnostics
    run = W.build(G, W.with_precision(W.lpa_cfg0(), precision), 5)
    for i in range(1, 4):
        run.make_step(i)
    dg = Diagnostics(run, diags=({"Type": "EnergyEM", "Step": 1},), out_folder=str(tmp_path))
    dat = dg.nrg_out({"Features": {"Return": True}})[0]
    sol = run.Solvers[0]
    ref = D.nrg_out(sol.Data["EG_fb"].numpy().astype(np.complex128),
                    D.energy_fact(sol.Args["idxM"], sol.Args["kr_g"], sol.Args["lengthR"], sol.Args["dkx"]))
    err = np.max(np.abs(dat - ref)) / np.max(np.abs(ref))
    print(f"case: nrg_out lpa_cfg0 {precision}: {err:.2e}", flush=True)
    assert err < 1e-12
    dg.do_diags(1)                                      # appends one line to nrg_0.txt
    assert np.allclose(np.loadtxt(tmp_path / "nrg_0.txt"), dat, rtol=1e-15, atol=0)


# ------------------------------------------------------------------ laser power per x (pwr_out)

def test_oracle_pwr_out_single_mode(port):
    """A field made of one Out row: fb_vec_out puts Out[k0, r, m0] * c * e^{i kx0 leftX} (a
    single kx, unnormalised backward FFT) on every x of row r, so the power is
    PwrFctr*2*dr * sum_r Rgrid[r] * |Out[k0,r,m0]|^2 * |c|^2 on every x."""
    nkx, nkr, M, dr = 16, 6, 2, 0.3
    r = np.random.default_rng(4)
    Out = np.asfortranarray(r.normal(size=(nkr, nkr, M)))
    Rg = dr * (np.arange(nkr + 1) - 0.5)
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.1)
    EG = np.zeros((nkx, nkr, M, 6), complex, order="F")
    EG[3, 2, 1, 1] = 2.0 - 1.0j
    p = D.pwr_out(port, EG, -1.3, kx, Out, Rg, dr)
    expect = D.PwrFctr * 2 * dr * (Rg[1:] * Out[2, :, 1] ** 2).sum() * 5.0
    assert np.allclose(p, expect, rtol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_pwr_out_matches_oracle(precision, port, tmp_path):
    import chimera_amd as G
    from chimera_amd import workloads as W
    from chimera_amd.diagnostics import Diagnostics
    run = W.build(G, W.with_precision(W.lpa_cfg0(), precision), 5)
    for i in range(1, 4):
        run.make_step(i)
    dg = Diagnostics(run, diags=({"Type": "Power", "Step": 1},), out_folder=str(tmp_path))
    p = dg.pwr_out({"Features": {"Return": True}})[0]
    sol = run.Solvers[0]
    A = sol.Args
    ref = D.pwr_out(port, sol.Data["EG_fb"].numpy().astype(np.complex128), float(A["leftX"]), A["kx"],
                    A["OutFull"].astype(sol.rt).astype(np.float64), A["RgridFull"], A["dr"])
    err = np.max(np.abs(p - ref)) / np.max(np.abs(ref))
    print(f"case: pwr_out lpa_cfg0 {precision}: {err:.2e}", flush=True)
    assert err < (1e-12 if precision == "fp64" else 2e-6)
    dg.do_diags(1)
    assert np.allclose(np.loadtxt(tmp_path / "pwrX_0.txt"), p, rtol=1e-15, atol=0)


# ------------------------------------------------------------------ dumps, density, phase space, spot

def test_oracle_intens_profo_closed_forms():
    """m = 0 alone: the same power at every angle, sum |F|^2; modes +-1 of
    equal amplitude a: |a e^{i th} + a e^{-i th}|^2 = 4 |a|^2 cos^2 th at the
    angles th = 2 pi (iO - 1)/(NO - 1)."""
    r = np.random.default_rng(21)
    nx, nr, NO = 5, 4, 13
    F0 = np.zeros((nx, nr + 1, 3, 3), complex)
    F0[:, :, 1, :] = r.normal(size=(nx, nr + 1, 3)) + 1j * r.normal(size=(nx, nr + 1, 3))
    P = D.intens_profo(F0, NO)
    expect = (np.abs(F0[:, 1:, 1, :]) ** 2).sum(axis=(0, 2))
    assert np.allclose(P, expect[None, :], rtol=1e-13)
    a = r.normal(size=(nx, nr + 1, 3)) + 1j * r.normal(size=(nx, nr + 1, 3))
    F1 = np.zeros((nx, nr + 1, 3, 3), complex)
    F1[:, :, 0, :] = a
    F1[:, :, 2, :] = a
    th = 2 * np.pi * np.arange(NO) / (NO - 1)
    expect = 4 * np.cos(th)[:, None] ** 2 * (np.abs(a[:, 1:]) ** 2).sum(axis=(0, 2))[None, :]
    assert np.allclose(D.intens_profo(F1, NO), expect, rtol=1e-12, atol=1e-12 * expect.max())


def test_oracle_phase_space_filter():
    """(1 + u^2) summed over the three components > 400, i.e. |u|^2 > 397."""
    u = np.array([[19.9, 0.0, 19.93, 0.0], [0.0, 0.0, 0.0, 12.0], [0.0, 0.0, 0.0, 15.0]])
    x = np.arange(12.0).reshape(3, 4)
    w = np.array([1.0, 2.0, 3.0, 4.0])
    pp = D.phase_space(x, u, w)
    assert pp.shape == (7, 1) and pp[6, 0] == 3.0 and np.array_equal(pp[:3, 0], x[:, 2])
    u[0, 0] = 19.95
    assert D.phase_space(x, u, w).shape == (7, 2)


def test_oracle_dens_on_grid_charge(port):
    """One particle of weight w at radius r0 in the middle of a cell: the m = 0
    density times the cell volumes 2 pi dx dr r sums to w (the ghost fold keeps
    the charge of the rows around the axis)."""
    dx, dr, Nx, Nr = 0.1, 0.3, 16, 8
    Rg = dr * (np.arange(Nr) - 0.5)
    leftX = -0.8
    x = np.array([[leftX + 7.5 * dx], [0.0], [2.2 * dr]])
    d = D.dens_on_grid(port, x, np.array([0.7]), leftX, Rg, dx, dr, Nx, 1)
    vol = 2 * np.pi * dx * dr * np.where(Rg > 0, Rg, 0.0)
    assert np.isclose((d[:, :, 0].real * vol[None, :]).sum(), 0.7, rtol=1e-13)


def _diag_run(precision, steps=3):
    import chimera_amd as G
    from chimera_amd import workloads as W
    run = W.build(G, W.with_precision(W.lpa_small(Nx=128, Nkr=32, nko=1), precision), 5)
    for i in range(1, steps + 1):
        run.make_step(i)
    return run


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_density_and_dumps(precision, port, tmp_path):
    from chimera_amd.diagnostics import Diagnostics
    run = _diag_run(precision)
    sol, el = run.Solvers[0], run.Particles[0]
    diags = ({"Type": "Fields", "Step": 1}, {"Type": "FieldsFB", "Step": 2}, {"Type": "Density", "Step": 1},
             {"Type": "Particles", "Step": 1})
    dg = Diagnostics(run, diags=diags, out_folder=str(tmp_path))
    A = el.Args
    for nko in (0, 1, 2):
        dens = el.get_dens_on_grid(nko)
        ref = D.dens_on_grid(port, el.host("coords").astype(float), el.host("weights").astype(float),
                             float(A["leftX"]), A["Rgrid"], A["dx"], A["dr"], A["Nx"], nko)
        err = np.linalg.norm(dens - ref) / np.linalg.norm(ref)
        print(f"case: get_dens_on_grid {precision} nko={nko}: {err:.2e}", flush=True)
        assert dens.shape == ref.shape and err < 1e-13
    f = dg.fld_out({"Features": {"Return": True}})[0]
    assert np.array_equal(f, sol.Data["EB"].numpy()[:, 1:])
    g = dg.fldfb_out({"Features": {"Return": True}})[0]
    assert np.array_equal(g, sol.Data["EG_fb"].numpy())
    dg.do_diags(1)
    assert np.array_equal(np.load(tmp_path / "ee_0_0000001.npy"), f)
    d1 = np.load(tmp_path / "dens_0_0000001.npy")
    assert d1.shape == (A["Nx"], A["Nr"] - 1, 2)
    assert (tmp_path / "phs0_0000001.npy").exists()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_phase_space_matches_oracle(precision):
    """A bunch with momenta spread over the filter's edge (|u| near 20): the
    selected particles, their order and their values."""
    from chimera_amd.diagnostics import Diagnostics
    run = _diag_run(precision, steps=1)
    el = run.Particles[0]
    r = np.random.default_rng(31)
    u = r.normal(size=(3, el.np)) * 12.0
    u[:, ::7] = [[19.93], [0.0], [0.0]]           # just above the edge (|u|^2 = 397.2)
    u[:, 3::11] = [[19.92], [0.0], [0.0]]         # just below (396.8)
    rt = el.Data["momenta"].dtype
    buf = el.Data["momenta"].numpy()
    buf[:el.np] = u.T.astype(rt)
    el.Data["momenta"].upload(buf)
    pp = Diagnostics(run).phs_out({"Features": {"Return": True}})[0]
    ref = D.phase_space(el.host("coords_halfstep").astype(float), el.host("momenta").astype(float),
                        el.host("weights").astype(float))
    print(f"case: phs_out {precision}: {pp.shape[1]} of {el.np} particles", flush=True)
    assert pp.shape == ref.shape and 0 < pp.shape[1] < el.np
    assert np.array_equal(pp, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_intens_profo_and_spot(precision, port):
    """intens_profO through the C ABI on a random field with modes -1..1
    against the oracle; then the 'Spot' profile of pwr_out on a single-mode
    run and get_spot_cartesian (the reference's griddata on it)."""
    import chimera_amd as G
    from chimera_amd import workloads as W
    from chimera_amd._lib import lib
    from chimera_amd.diagnostics import Diagnostics
    r = np.random.default_rng(41)
    nkx, nr, nm, NO = 48, 12, 3, 60
    F = np.asfortranarray(r.normal(size=(nkx, nr + 1, nm, 3)) + 1j * r.normal(size=(nkx, nr + 1, nm, 3)))
    out = np.zeros((NO, nr), order="F")
    if precision == "fp64":
        st = lib().chimera_intens_profo(out.ctypes.data, F.ctypes.data, NO, nkx, nr, nm)
        ref = D.intens_profo(F, NO)
    else:
        F8 = np.asfortranarray(F.astype(np.complex64))
        st = lib().chimera_intens_profo_c8(out.ctypes.data, F8.ctypes.data, NO, nkx, nr, nm)
        ref = D.intens_profo(F8.astype(np.complex128), NO)
    assert st == 0
    err = np.max(np.abs(out - ref)) / np.max(np.abs(ref))
    print(f"case: intens_profo {precision}: {err:.2e}", flush=True)
    assert err < 1e-13
    assert lib().chimera_intens_profo(out.ctypes.data, F.ctypes.data, NO, nkx, nr, 2) != 0   # even mode count
    run = W.build(G, W.with_precision(W.lpa_small(Nx=128, Nkr=32, nko=0), precision), 5)
    run.make_step(1)
    dg = Diagnostics(run)
    pw, prof = dg.pwr_out({"Features": {"Return": True, "Spot": True}})[0]
    sol = run.Solvers[0]
    A = sol.Args
    e = np.asfortranarray(sol.Data["EG_fb"].numpy().astype(np.complex128))
    eb = port.fb_eb_out(np.zeros((A["Nx"], A["OutFull"].shape[1] + 1, 1, 6), complex, order="F"), e,
                        np.asfortranarray(e[..., :3]), float(A["leftX"]), A["kx"],
                        A["OutFull"].astype(sol.rt).astype(np.float64))
    ref = D.intens_profo(np.asfortranarray(eb[..., :3]), NO)
    err = np.max(np.abs(prof - ref)) / np.max(np.abs(ref))
    print(f"case: pwr_out Spot {precision}: {err:.2e}", flush=True)
    assert err < (1e-12 if precision == "fp64" else 2e-6)
    val, ext = dg.get_spot_cartesian(bins=(40, 40))
    assert val.shape == (40, 40) and np.all(np.isfinite(val)) and ext[1] == A["RgridFull"].max()
