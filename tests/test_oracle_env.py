"""Pins the oracle's restatement of the envelope (KxShift) routines of
grid_deps_env.f90 and fb_math_env.f90 (cpu_ref.c) against the reference's own
Fortran, compiled from the reference sources by oracle/Makefile: bit-identical
results (both sides plain IEEE double arithmetic in the same order), for mode
axes -nkO..nkO with nkO = 0, 1, 2, and for the complex PSATD tables of the
envelope solver through maxwell_push_wo_spchrg.

Then what the mathematics fixes, so that a mistake made the same way in the
restatement and in a reading of the reference would still show:
* the envelope deposit of one particle with kx0 = 0 and nkO = 0 is the plain
  deposit's m = 0 z-current (the same weights), and the envelope density is
  the plain density times the weight (the weight applied twice);
* a carrier kx0 multiplies the deposits by e^{-i kx0 x} (w e^{-i kx0 x} twice
  for the density) and the gather by e^{+i kx0 x};
* gathering the field made of one node value gives the interpolation weight
  times the carrier and phase, mode by mode.
"""
import numpy as np
import pytest

from tests._cases import C, F, Grid, particles, rng


def _env_grid(nko):
    return Grid(nx=36, nkr=10, nko=nko)


@pytest.mark.parametrize("nko", [0, 1, 2])
def test_env_deposit_gather_match_reference(port, reference, nko):
    g = _env_grid(nko)
    coord, mom, w = particles(g, 2500, seed=40 + nko)
    nm = 2 * nko + 1
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr, 2.7)
    z = np.zeros((g.nx, g.nkr + 1, nm, 3), complex, order="F")
    a = port.dep_curr_env(coord, mom, w, z.copy(order="F"), *args)
    b = reference.dep_curr_env(coord, mom, w, z.copy(order="F"), *args)
    assert np.array_equal(a, b) and np.any(a[..., 2]) and not np.any(a[..., :2])
    zs = np.zeros((g.nx, g.nkr + 1, nm), complex, order="F")
    assert np.array_equal(port.dep_dens_env(coord, w, zs.copy(order="F"), *args),
                          reference.dep_dens_env(coord, w, zs.copy(order="F"), *args))
    EB = C(rng(7 + nko), g.nx, g.nkr + 1, nm, 6)
    ft = F(rng(8), 6, coord.shape[1])
    assert np.array_equal(port.proj_fld_env(coord, w, EB, ft.copy(order="F"), *args),
                          reference.proj_fld_env(coord, w, EB, ft.copy(order="F"), *args))
    assert np.array_equal(port.eb_correction_env(EB.copy(order="F")), reference.eb_correction_env(EB.copy(order="F")))


@pytest.mark.parametrize("nko", [0, 1, 2])
def test_env_fb_math_matches_reference(port, reference, nko):
    r = rng(60 + nko)
    nkx, nkr = 24, 9
    nm = 2 * nko + 1
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.1) + 3.0
    Dp, Dm = F(r, nkr, nkr, nm + 2), F(r, nkr, nkr, nm + 2)
    v, s = C(r, nkx, nkr, nm, 3), C(r, nkx, nkr, nm)
    z3 = np.zeros((nkx, nkr, nm, 3), complex, order="F")
    z1 = np.zeros((nkx, nkr, nm), complex, order="F")
    for name, out, inp in (("fb_rot_env", z3, v), ("fb_grad_env", z3, s), ("fb_div_env", z1, v)):
        a = getattr(port, name)(out.copy(order="F"), inp, Dp, Dm, kx)
        b = getattr(reference, name)(out.copy(order="F"), inp, Dp, Dm, kx)
        assert np.array_equal(a, b), name
    assert np.array_equal(port.fb_graddiv_env(v.copy(order="F"), Dp, Dm, kx),
                          reference.fb_graddiv_env(v.copy(order="F"), Dp, Dm, kx))


def test_env_maxwell_complex_tables_match_reference(port, reference):
    r = rng(70)
    sh = (20, 8, 3)
    EG, j = C(r, *sh, 6), C(r, *sh, 3)
    C1, C2 = C(r, *sh, 3), C(r, *sh, 3)
    assert np.array_equal(port.maxwell_push_wo_spchrg(EG.copy(order="F"), j, C1, C2),
                          reference.maxwell_push_wo_spchrg(EG.copy(order="F"), j, C1, C2))


def test_env_deposits_reduce_to_the_plain_ones(port):
    """kx0 = 0, nkO = 0: the envelope z-current equals the plain deposit's
    m = 0 z-current, the envelope density the plain density of the squared
    weights; with a carrier kx0, one particle's deposits carry e^{-i kx0 x}
    (density: twice)."""
    g = Grid(nx=30, nkr=8, nko=0)
    coord, mom, w = particles(g, 1500, seed=91)
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr)
    z = np.zeros((g.nx, g.nkr + 1, 1, 3), complex, order="F")
    je = port.dep_curr_env(coord, mom, w, z.copy(order="F"), *args, 0.0)
    jp = port.dep_curr(coord, mom, w, z.copy(order="F"), *args)
    assert np.allclose(je[..., 2], jp[..., 2], rtol=1e-13, atol=1e-15 * np.abs(jp).max())
    zs = np.zeros((g.nx, g.nkr + 1, 1), complex, order="F")
    de = port.dep_dens_env(coord, w, zs.copy(order="F"), *args, 0.0)
    dp = port.dep_dens(coord, w ** 2, zs.copy(order="F"), *args)
    assert np.allclose(de, dp, rtol=1e-13, atol=1e-15 * np.abs(dp).max())
    one = (np.asfortranarray(coord[:, 5:6]), np.asfortranarray(mom[:, 5:6]), w[5:6].copy())
    kx0, x = 1.3, coord[0, 5]
    a = port.dep_curr_env(*one, z.copy(order="F"), *args, kx0)
    b = port.dep_curr_env(*one, z.copy(order="F"), *args, 0.0)
    assert np.allclose(a, b * np.exp(-1j * kx0 * x), rtol=1e-13, atol=1e-16)
    a = port.dep_dens_env(one[0], one[2], zs.copy(order="F"), *args, kx0)
    b = port.dep_dens_env(one[0], one[2], zs.copy(order="F"), *args, 0.0)
    assert np.allclose(a, b * np.exp(-2j * kx0 * x), rtol=1e-13, atol=1e-16)


@pytest.mark.parametrize("nko", [0, 1])
def test_env_gather_of_one_node(port, nko):
    """A field of one node (node (ix+a, ir+b) of mode m, component l) gathers
    at a particle to Re(S_a S_b e^{i kx0 x} phase_m F) with phase_m =
    ((y + i z)/r)^|m|, conjugated for m < 0."""
    g = Grid(nx=20, nkr=6, nko=nko)
    nm = 2 * nko + 1
    x, y, zc = g.leftX + 7.3 * g.dx, 0.9, -0.6
    r = np.hypot(y, zc)
    coord = np.asfortranarray([[x], [y], [zc]])
    fx = (x - g.leftX) / g.dx
    ix = int(np.floor(fx))
    ir = int(np.floor((r - g.Rgrid[0]) / g.dr))
    Sx = [1 - (fx - ix), fx - ix]
    Sr1 = (r - g.Rgrid[ir]) / g.dr
    Sr = [1 - Sr1, Sr1]
    kx0 = 0.8
    for m in range(-nko, nko + 1):
        for a, b in ((0, 0), (1, 1)):
            F_ = np.zeros((g.nx, g.nkr + 1, nm, 6), complex, order="F")
            val = 0.7 - 0.4j
            F_[ix + a, ir + b, m + nko, 4] = val
            out = port.proj_fld_env(coord, np.ones(1), F_, np.zeros((6, 1), order="F"), g.leftX, g.Rgrid,
                                    1 / g.dx, 1 / g.dr, kx0)
            ph = ((y + 1j * zc) / r) ** abs(m)
            ph = np.conj(ph) if m < 0 else ph
            expect = (Sx[a] * Sr[b] * np.exp(1j * kx0 * x) * ph * val).real
            assert np.isclose(out[4, 0], expect, rtol=1e-13, atol=1e-15), (m, a, b)
            assert not np.any(out[[0, 1, 2, 3, 5], 0])
