"""Pins the oracle (oracle/cpu_ref.c) before anything is compared with it.

* Against the reference itself: the reference's own Fortran for every routine
  that builds without FFTW (oracle/_ref, compiled from the reference sources by
  oracle/Makefile) -- results must be bit-identical, since both sides are plain
  IEEE double arithmetic in the same order.
* For fb_io.f90 (FFTW3 absent, so not buildable): against numpy.fft and dense
  linear algebra, and against closed forms (a single Bessel mode maps to one FB
  coefficient; a pure phase shift; the absorbing profile on a constant line).
* Invariants that a dropped term, wrong sign or transposed index would break:
  deposition conserves charge, gather is the adjoint of deposition, Boris
  conserves |u| without E and rotates by 2 atan|t|.  The operator tables the
  FB operators take (D+, D-) are pinned by the Bessel recurrences in
  tests/test_oracle_driver.py.
"""
import numpy as np
import pytest

from tests._cases import C, F, Grid, particles, rel_l2, rel_max, rng


# ------------------------------------------------------------------ oracle == reference Fortran (bit-exact)

def test_push_matches_reference(port, reference):
    r = rng(3)
    u, f = F(r, 3, 2000), F(r, 6, 2000)
    assert np.array_equal(port.push_velocs(u.copy(order="F"), f, 0.37), reference.push_velocs(u.copy(order="F"), f, 0.37))
    x, xc = F(r, 3, 2000), F(r, 3, 2000)
    a = port.push_coords(x.copy(order="F"), u, xc.copy(order="F"), 0.05)
    b = reference.push_coords(x.copy(order="F"), u, xc.copy(order="F"), 0.05)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("nko", [0, 1, 3])
def test_deposit_gather_match_reference(port, reference, nko):
    g = Grid(nx=40, nkr=12, nko=nko)
    coord, mom, w = particles(g, 3000, seed=10 + nko)
    z = np.zeros((g.nx, g.nkr + 1, nko + 1, 3), complex, order="F")
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr)
    assert np.array_equal(port.dep_curr(coord, mom, w, z.copy(order="F"), *args),
                          reference.dep_curr(coord, mom, w, z.copy(order="F"), *args))
    zs = np.zeros((g.nx, g.nkr + 1, nko + 1), complex, order="F")
    assert np.array_equal(port.dep_dens(coord, w, zs.copy(order="F"), *args),
                          reference.dep_dens(coord, w, zs.copy(order="F"), *args))
    EB = C(rng(5), g.nx, g.nkr + 1, nko + 1, 6)
    ft = F(rng(6), 6, coord.shape[1])
    assert np.array_equal(port.proj_fld(coord, w, EB, ft.copy(order="F"), *args),
                          reference.proj_fld(coord, w, EB, ft.copy(order="F"), *args))
    assert np.array_equal(port.eb_correction(EB.copy(order="F")), reference.eb_correction(EB.copy(order="F")))


@pytest.mark.parametrize("nko", [1, 2, 3])
def test_fb_math_matches_reference(port, reference, nko):
    r = rng(20 + nko)
    nkx, nkr = 32, 10
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.1)
    Dp, Dm = F(r, nkr, nkr, nko + 2), F(r, nkr, nkr, nko + 2)
    v, s = C(r, nkx, nkr, nko + 1, 3), C(r, nkx, nkr, nko + 1)
    z3 = np.zeros((nkx, nkr, nko + 1, 3), complex, order="F")
    z1 = np.zeros((nkx, nkr, nko + 1), complex, order="F")
    for name, out, inp in (("fb_rot", z3, v), ("fb_grad", z3, s), ("fb_div", z1, v)):
        a = getattr(port, name)(out.copy(order="F"), inp, Dp, Dm, kx)
        b = getattr(reference, name)(out.copy(order="F"), inp, Dp, Dm, kx)
        assert np.array_equal(a, b), name
    assert np.array_equal(port.fb_graddiv(v.copy(order="F"), Dp, Dm, kx),
                          reference.fb_graddiv(v.copy(order="F"), Dp, Dm, kx))


def test_maxwell_match_reference(port, reference):
    r = rng(30)
    sh = (24, 10, 3)
    EG = C(r, *sh, 6)
    j, gn, gp, gd = C(r, *sh, 3), C(r, *sh, 3), C(r, *sh, 3), C(r, *sh, 3)
    C1, C2 = F(r, *sh, 5), F(r, *sh, 5)
    assert np.array_equal(port.maxwell_push_with_spchrg(EG.copy(order="F"), j, gn, gp, C1, C2),
                          reference.maxwell_push_with_spchrg(EG.copy(order="F"), j, gn, gp, C1, C2))
    assert np.array_equal(port.maxwell_push_wo_spchrg(EG.copy(order="F"), j, C1[..., :3], C2[..., :3]),
                          reference.maxwell_push_wo_spchrg(EG.copy(order="F"), j, C1[..., :3], C2[..., :3]))
    CC1, CC2 = C(r, *sh, 2), C(r, *sh, 2)
    assert np.array_equal(port.maxwell_init_push(EG.copy(order="F"), j, gn, CC1, CC2),
                          reference.maxwell_init_push(EG.copy(order="F"), j, gn, CC1, CC2))
    w2 = F(r, *sh)
    assert np.array_equal(port.poiss_corr(j.copy(order="F"), gd, gn, gp, 20.0, w2),
                          reference.poiss_corr(j.copy(order="F"), gd, gn, gp, 20.0, w2))
    DT = C(r, sh[0])
    assert np.array_equal(port.poiss_corr_stat(j.copy(order="F"), gd, gn, DT, w2),
                          reference.poiss_corr_stat(j.copy(order="F"), gd, gn, DT, w2))
    kx = 2 * np.pi * np.fft.fftfreq(sh[0], 0.07)
    assert np.array_equal(port.field_drift(EG.copy(order="F"), kx, 0.999, 0.035),
                          reference.field_drift(EG.copy(order="F"), kx, 0.999, 0.035))
    for name, A in (("omp_mult_vec", w2), ("omp_add_vec", gn)):
        assert np.array_equal(getattr(port, name)(j.copy(order="F"), A), getattr(reference, name)(j.copy(order="F"), A))
    s = C(r, *sh)
    assert np.array_equal(port.omp_mult_scl(s.copy(order="F"), w2), reference.omp_mult_scl(s.copy(order="F"), w2))
    assert np.array_equal(port.omp_add_scl(s.copy(order="F"), s), reference.omp_add_scl(s.copy(order="F"), s))


def _mapped_case(seed, tap):
    """Particles over and around a field map of 40 nodes (x0 = -1.2, dx = 0.1),
    none in [x0 + dx, x0 + 1.5 dx) where the reference reads the node before
    the map; a seeded map; params of Undul_Mapped(_Tap)."""
    r = rng(seed)
    nx, x0, dx = 40, -1.2, 0.1
    x = r.uniform(x0 - 0.3, x0 + (nx + 3) * dx, 3000)
    x = x[(x < x0 + dx) | (x >= x0 + 1.5 * dx)]
    coord = np.asfortranarray(np.vstack([x, r.normal(0, 0.3, x.size), r.normal(0, 0.3, x.size)]))
    a0 = F(r, 2, nx)
    params = np.array([0.9, x0, dx, 3.1, 0.25] if tap else [0.9, x0, dx])
    return coord, a0, params, F(r, 6, x.size)


@pytest.mark.parametrize("tap", [0, 1])
def test_undul_mapped_matches_reference(port, reference, tap):
    name = "undul_mapped_tap" if tap else "undul_mapped"
    coord, a0, params, ft = _mapped_case(60 + tap, tap)
    a = getattr(port, name)(coord, ft.copy(order="F"), 0.0, a0, params)
    b = getattr(reference, name)(coord, ft.copy(order="F"), 0.0, a0, params)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, ft)


def test_wave_devices_match_reference(port, reference):
    """PlaneWave and GaussBeam (devices.f90:205-297), and the closed forms they
    reduce to: on the flat top of the plane wave Fld(3) gains a0 sin(phase);
    at the centre of the Gaussian beam the field is a0 sin(0) = 0 and one
    quarter wavelength ahead it is a0 exp(-(lambda/4)^2/Lx^2)."""
    r = rng(63)
    n = 3000
    coord = np.asfortranarray(np.vstack([r.uniform(-1, 9, n), r.normal(0, 1, n), r.normal(0, 1, n)]))
    ft = F(r, 6, n)
    pw = np.array([0.8, 1.3, 0.5, 7.0, 1.1, 0.35, 0.2])
    a = port.planewave(coord, ft.copy(order="F"), 0.7, pw)
    assert np.array_equal(a, reference.planewave(coord, ft.copy(order="F"), 0.7, pw))
    gb = np.array([1.3, 0.9, 3.0, 0.2, -0.1, 1.5, 0.8, 1.1])
    b = port.gaussbeam(coord, ft.copy(order="F"), 0.4, 1.7, gb)
    assert np.array_equal(b, reference.gaussbeam(coord, ft.copy(order="F"), 0.4, 1.7, gb))
    z = np.zeros((6, n), order="F")
    x, y = coord[0], coord[1]
    flat = (x >= 0.5 + 1.1) & (x <= 0.5 + 7.0 - 1.1)
    ph = 2 * np.pi / 1.3 * (x * np.cos(0.35) + y * np.sin(0.35) - 0.7) + 0.2
    f = port.planewave(coord, z.copy(order="F"), 0.7, pw)
    np.testing.assert_allclose(f[2, flat], 0.8 * np.sin(ph[flat]), atol=1e-13)
    np.testing.assert_allclose(f[4, flat], -np.cos(0.35) * f[2, flat], rtol=1e-15)
    assert not f[:, (x <= 0.5) | (x >= 7.5)].any()
    c = np.asfortranarray(np.array([[3.0 + 0.9 * 0.4, 3.0 + 0.9 * 0.4 + 1.3 / 4], [0.2, 0.2], [-0.1, -0.1]]))
    g = port.gaussbeam(c, np.zeros((6, 2), order="F"), 0.4, 1.7, gb)
    assert abs(g[2, 0]) < 1e-14          # xp is zero up to the rounding of the centre
    np.testing.assert_allclose(g[2, 1], 1.7 * np.exp(-(1.3 / 4) ** 2 / 1.5 ** 2), rtol=1e-14)
    np.testing.assert_allclose(g[4], -0.9 * g[2], rtol=1e-15)


def test_undul_mapped_closed_form(port):
    """A constant map c: the quadratic weights sum to one, so inside the map
    Fld(5) gains c cosh(ku y) and Fld(4) gains c sinh(ku y); outside, nothing.
    Tapered, the gain is scaled by the linear taper."""
    r = rng(62)
    nx, x0, dx, lam = 50, 0.0, 0.2, 1.3
    x = r.uniform(-1.0, nx * dx + 1.0, 2000)
    y = r.normal(0, 0.4, x.size)
    coord = np.asfortranarray(np.vstack([x, y, 0 * x]))
    a0 = np.asfortranarray(np.vstack([np.full(nx, 2.5), np.full(nx, -0.7)]))
    z = np.zeros((6, x.size), order="F")
    inside = (x >= x0 + 1.5 * dx) & (x <= x0 + (nx - 1) * dx)
    outside = (x < x0 + dx) | (x > x0 + (nx - 1) * dx)
    ku = 2 * np.pi / lam
    f = port.undul_mapped(coord, z.copy(order="F"), 0.0, a0, np.array([lam, x0, dx]))
    np.testing.assert_allclose(f[4, inside], 2.5 * np.cosh(ku * y[inside]), rtol=1e-14)
    np.testing.assert_allclose(f[3, inside], -0.7 * np.sinh(ku * y[inside]), rtol=1e-13, atol=1e-15)
    assert not f[:, outside].any()
    Lx, taper = 8.0, 0.3
    amp = 1 + taper * (x - 0.5 * (nx * dx - Lx) - x0 - 0.5 * Lx) / (0.5 * Lx)
    f = port.undul_mapped_tap(coord, z.copy(order="F"), 0.0, a0, np.array([lam, x0, dx, Lx, taper]))
    np.testing.assert_allclose(f[4, inside], amp[inside] * 2.5 * np.cosh(ku * y[inside]), rtol=1e-13)


def test_particle_tools_match_reference(port, reference):
    g = Grid(nx=30, nkr=10, nko=1)
    coord, mom, w = particles(g, 2500, seed=40)
    p = np.array([1.7, 0.9, -2.0, 3.5])
    ft = F(rng(41), 6, coord.shape[1])
    assert np.array_equal(port.undul_analytic(coord, ft.copy(order="F"), 0.0, p),
                          reference.undul_analytic(coord, ft.copy(order="F"), 0.0, p))
    pt = np.array([1.7, 0.9, -2.0, 3.5, 0.3])
    a = port.undul_analytic_taper(coord, ft.copy(order="F"), 0.0, pt)
    assert np.array_equal(a, reference.undul_analytic_taper(coord, ft.copy(order="F"), 0.0, pt))
    assert not np.array_equal(a, port.undul_analytic(coord, ft.copy(order="F"), 0.0, pt[:4]))
    # no taper: the untapered undulator, bit for bit
    assert np.array_equal(port.undul_analytic_taper(coord, ft.copy(order="F"), 0.0, np.append(p, 0.0)),
                          port.undul_analytic(coord, ft.copy(order="F"), 0.0, p))
    Xg = g.leftX + g.dx * np.arange(g.nx)
    RP = rng(42).random((g.nx, g.nkr + 1))
    pX, pR, pO = np.mgrid[1:2:2j, 1:2:2j, 1:4:4j]
    PX, PR = (pX.ravel() - 0.5) / 2, (pR.ravel() - 0.5) / 2
    PO = np.exp(2j * np.pi * (pO.ravel() - 1) / 4)
    co = np.zeros((4, g.nx * (g.nkr + 1) * 16), order="F")
    a = port.genparts(co.copy(order="F"), Xg, g.Rgrid, RP, PX, PR, PO)
    b = reference.genparts(co.copy(order="F"), Xg, g.Rgrid, RP, PX, PR, PO)
    assert a[1] == b[1] == (g.nx - 1) * g.nkr * 16
    assert np.array_equal(a[0], b[0])
    lims = np.array([g.leftX + 0.5, g.leftX + 2.5, 0.0, 4.0])
    a, b = port.sortpartsout(coord, lims), reference.sortpartsout(coord, lims)
    assert a[1] == b[1] and np.array_equal(a[0], b[0])
    Xg0 = g.leftX + g.dx * np.arange(g.nx + 1)
    a, b = port.chunk_coords_boundaries(coord, lims, Xg0, 4), reference.chunk_coords_boundaries(coord, lims, Xg0, 4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    idx = rng(43).permutation(coord.shape[1])[:1700]
    assert np.array_equal(port.align_data_vec(coord.copy(order="F"), idx), reference.align_data_vec(coord.copy(order="F"), idx))
    assert np.array_equal(port.align_data_scl(w.copy(), idx), reference.align_data_scl(w.copy(), idx))


# ------------------------------------------------------------------ fb_io.f90: numpy and closed forms

def _np_fb_in(vec, leftX, kx, In):
    # DHT over rows 1..nr, numpy FFT along x, phase exp(-i kx leftX)  (fb_io.f90:42-53)
    A = np.einsum("xrml,rkm->xkml", vec[:, 1:], In)
    return np.fft.fft(A, axis=0) * np.exp(-1j * kx * leftX)[:, None, None, None]


def test_fb_in_matches_numpy(port):
    r = rng(50)
    nkx, nr, nkr, M = 64, 12, 11, 3
    vec = C(r, nkx, nr + 1, M, 3)
    In = F(r, nr, nkr, M)
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.07)
    out = port.fb_vec_in(np.zeros((nkx, nkr, M, 3), complex, order="F"), vec, -1.3, kx, In)
    assert rel_l2(out, _np_fb_in(vec, -1.3, kx, In)) < 1e-14
    outs = port.fb_scl_in(np.zeros((nkx, nkr, M), complex, order="F"), vec[..., 0].copy(order="F"), -1.3, kx, In)
    assert rel_l2(outs, _np_fb_in(vec[..., :1], -1.3, kx, In)[..., 0]) < 1e-14


def test_fb_in_non_power_of_two_line(port):
    r = rng(51)
    nkx, nr, nkr, M = 24, 5, 5, 2
    vec = C(r, nkx, nr + 1, M, 3)
    In = F(r, nr, nkr, M)
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.2)
    out = port.fb_vec_in(np.zeros((nkx, nkr, M, 3), complex, order="F"), vec, 0.4, kx, In)
    assert rel_l2(out, _np_fb_in(vec, 0.4, kx, In)) < 1e-13


def test_fb_eb_out_matches_numpy(port):
    r = rng(52)
    nkx, nr, nkr, M = 32, 9, 9, 2
    e, b = C(r, nkx, nkr, M, 6), C(r, nkx, nkr, M, 3)
    Out = F(r, nkr, nr, M)
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.1)
    eb = port.fb_eb_out(np.ones((nkx, nr + 1, M, 6), complex, order="F"), e, b, 0.77, kx, Out)
    ph = np.exp(1j * kx * 0.77)[:, None, None, None]
    ref_e = np.fft.ifft(np.einsum("xkml,krm->xrml", e[..., :3], Out) * ph, axis=0) * nkx
    ref_b = np.fft.ifft(np.einsum("xkml,krm->xrml", b, Out) * ph, axis=0) * nkx
    assert rel_l2(eb[:, 1:, :, :3], ref_e) < 1e-14
    assert rel_l2(eb[:, 1:, :, 3:], ref_b) < 1e-14
    assert not np.any(eb[:, 0])                      # ghost row zeroed (fb_io.f90:199)


def test_fb_single_bessel_mode(port):
    """J_m(kr_j r) exp(i kx_q x) on the grid maps to exactly one FB coefficient
    when In = inv(Out) (solvers.py:742-743): the transform is diagonal."""
    from scipy.special import jn, jn_zeros
    nkx, nkr, M, dx, dr = 32, 16, 2, 0.1, 0.3
    lengthR = dr * (nkr + 0.5)
    Rg = dr * (np.arange(nkr + 1) - 0.5)
    kx = 2 * np.pi * np.fft.fftfreq(nkx, dx)
    In = np.zeros((nkr, nkr, M), order="F")
    for m in range(M):
        kr = jn_zeros(m, nkr) / lengthR
        Out = jn(m, Rg[1:, None] * kr[None, :])        # (r, k)
        In[:, :, m] = np.linalg.inv(Out).T                 # swapped axes (solvers.py:152-153): In(r, k)
    q, j, m = 5, 3, 1
    x = np.arange(nkx) * dx
    kr = jn_zeros(m, nkr) / lengthR
    vec = np.zeros((nkx, nkr + 1, M), complex, order="F")
    vec[:, 1:, m] = np.exp(1j * kx[q] * x)[:, None] * jn(m, Rg[1:] * kr[j])[None, :]
    out = port.fb_scl_in(np.zeros((nkx, nkr, M), complex, order="F"), vec, 0.0, kx, In)
    expect = np.zeros_like(out)
    expect[q, j, m] = nkx
    assert np.abs(out - expect).max() < 1e-9 * nkx


def test_fb_filtr_closed_forms(port):
    nkx, nkr, M = 64, 4, 2
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.1)
    filt = np.linspace(0.0, 1.0, 20)
    # leftX = 0, a line that is constant in x: the filtered line is the profile
    v = np.zeros((nkx, nkr, M, 3), complex, order="F")
    v[0] = 3.0 * nkx          # FFT of a constant 3 (unnormalised)
    out = port.fb_filtr(v.copy(order="F"), 0.0, kx, filt, 0)
    line = np.fft.ifft(out[:, 0, 0, 0])
    prof = np.ones(nkx) * 3.0
    prof[:20] *= filt
    assert np.abs(line - prof).max() < 1e-12
    # with a profile of ones the filter is the identity for any leftX
    r = rng(53)
    w = C(r, nkx, nkr, M, 3)
    out = port.fb_filtr(w.copy(order="F"), 1.234, kx, np.ones(30), 0)
    assert rel_max(out, w) < 1e-13
    with pytest.raises(ValueError):
        port.fb_filtr(w.copy(order="F"), 0.0, kx, filt, 1)


# ------------------------------------------------------------------ invariants of the particle kernels

def test_deposit_conserves_charge_and_gather_is_adjoint(port):
    g = Grid(nx=48, nkr=12, nko=2)
    coord, mom, w = particles(g, 3000, seed=60)
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr)
    dens = port.dep_dens(coord, w, np.zeros((g.nx, g.nkr + 1, 3), complex, order="F"), *args)
    # m = 0: the four bilinear weights sum to 1; the ghost fold moves row 0 into row 1 with a minus sign,
    # so total charge minus twice the ghost-row share is conserved: sum = sum(w) - 2*sum(ghost part)
    rp = np.hypot(coord[1], coord[2])
    fr = (rp - g.Rgrid[0]) / g.dr
    ir = np.floor(fr).astype(int)
    s1 = (rp - g.Rgrid[ir]) / g.dr
    ghost = np.where(ir == 0, (1 - s1) * w, 0.0)
    assert abs(dens[:, :, 0].sum().real - (w.sum() - 2 * ghost.sum())) < 1e-11 * w.sum()
    assert abs(dens[:, :, 0].sum().imag) < 1e-12 * w.sum()
    # adjointness: with phases e^{+im phi} in the gather and e^{-im phi} in the deposit,
    # sum_p w_p Fgathered_p = Re sum_nodes F * conj(rho).  Rows 0 and 1 of F are zeroed so
    # that the ghost fold (which only touches those rows) drops out.
    r = rng(61)
    Fld = C(r, g.nx, g.nkr + 1, 3, 6)
    Fld[:, :2] = 0.0
    ft = port.proj_fld(coord, w, Fld, np.zeros((6, coord.shape[1]), order="F"), *args)
    for l in (0, 4):
        lhs = (ft[l] * w).sum()
        rhs = np.real((Fld[..., l] * np.conj(dens)).sum())
        assert abs(lhs - rhs) < 1e-12 * np.abs(ft[l] * w).sum()


def test_boris_closed_forms(port):
    r = rng(70)
    n = 500
    u = F(r, 3, n)
    # E = 0: |u| conserved
    fld = np.zeros((6, n), order="F")
    fld[3:] = F(r, 3, n)
    out = port.push_velocs(u.copy(order="F"), fld, 0.2)
    assert np.allclose(np.linalg.norm(out, axis=0), np.linalg.norm(u, axis=0), rtol=1e-14, atol=0)
    # uniform B along x: rotation angle 2*atan(|t|), t = dt/2 * B / gamma
    B = 1.3
    fld = np.zeros((6, n), order="F")
    fld[3] = B
    out = port.push_velocs(u.copy(order="F"), fld, 0.2)
    gam = np.sqrt(1 + (u ** 2).sum(0))
    ang = 2 * np.arctan(0.1 * B / gam)
    a0 = np.arctan2(u[2], u[1])
    a1 = np.arctan2(out[2], out[1])
    assert np.allclose(np.angle(np.exp(1j * (a0 - a1))), ang, atol=1e-12)
    # pure E kick: du = dt*E
    fld = np.zeros((6, n), order="F")
    fld[:3] = F(r, 3, n)
    out = port.push_velocs(u.copy(order="F"), fld, 0.2)
    assert np.allclose(out - u, 0.2 * fld[:3], rtol=0, atol=1e-14)


def test_particles_outside_the_grid_are_skipped(port):
    """A particle whose cell lies outside the arrays (x before leftX, x at or
    past the last node, r at or past Rgrid[nr]) deposits nothing and gathers
    zero fields; the reference has no x bound there (grid_deps.f90:46-47)."""
    g = Grid(nx=40, nkr=10, nko=1)
    coord, mom, w = particles(g, 400, seed=62)
    xmax = g.leftX + (g.nx - 1) * g.dx
    out = np.asfortranarray(np.array([[g.leftX - 0.3 * g.dx, 0.2, 0.1], [xmax, 0.2, 0.1],
                                      [xmax + 0.7 * g.dx, 0.1, 0.3], [g.leftX + 5 * g.dx, 0.0, g.Rgrid[-1]]]).T)
    mo = np.asfortranarray(np.ones((3, 4)))
    wo = np.ones(4)
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr)
    mix = lambda a, b: np.asfortranarray(np.concatenate([a, b], axis=-1))
    z = lambda *s: np.zeros(s, complex, order="F")
    for fn, extra, shape in (("dep_curr", (mom, mo), (g.nx, g.nkr + 1, 2, 3)), ("dep_dens", None, (g.nx, g.nkr + 1, 2))):
        f = getattr(port, fn)
        if extra:
            a = f(coord, mom, w, z(*shape), *args)
            b = f(mix(coord, out), mix(mom, mo), mix(w, wo), z(*shape), *args)
        else:
            a = f(coord, w, z(*shape), *args)
            b = f(mix(coord, out), mix(w, wo), z(*shape), *args)
        assert np.array_equal(a, b), fn
    r = rng(63)
    Fld = C(r, g.nx, g.nkr + 1, 2, 6)
    ft = port.proj_fld(mix(coord, out), mix(w, wo), Fld, np.zeros((6, coord.shape[1] + 4), order="F"), *args)
    assert np.all(ft[:, -4:] == 0.0) and np.any(ft[:, :-4] != 0.0)


def test_fma_variant_rounds_differently_only(port):
    """The port compiled with FMA contraction (the oracle's rounding-floor
    variant) computes the same arithmetic: it agrees with the plain port to
    rounding (1e-14 relative on a deposit and a Hankel transform) and is not
    bit-identical to it (the contraction is on)."""
    from oracle.fimera_ref import Fimera
    fma = Fimera("port_fma")
    g = Grid(nx=48, nkr=12, nko=2)
    coord, mom, w = particles(g, 3000, seed=77)
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr)
    z = np.zeros((g.nx, g.nkr + 1, 3, 3), complex, order="F")
    a = port.dep_curr(coord, mom, w, z.copy(order="F"), *args)
    b = fma.dep_curr(coord, mom, w, z.copy(order="F"), *args)
    assert np.linalg.norm(a - b) <= 1e-14 * np.linalg.norm(a)
    r = rng(78)
    vec, In = C(r, g.nx, g.nkr + 1, 3, 3), F(r, g.nkr, g.nkr, 3)
    zf = np.zeros((g.nx, g.nkr, 3, 3), complex, order="F")
    a = port.fb_vec_in(zf.copy(order="F"), vec, g.leftX, g.kx, In)
    b = fma.fb_vec_in(zf.copy(order="F"), vec, g.leftX, g.kx, In)
    assert np.linalg.norm(a - b) <= 1e-14 * np.linalg.norm(a)
    assert not np.array_equal(a, b)
