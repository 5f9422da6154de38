"""GPU parity of every drop-in entry point against the oracle, through the C ABI.

Floating-point results must agree with the FP64 oracle to a relative error of
1e-12 (L2 over the output array) or tighter where the operation is elementwise;
the differences come only from fused multiply-adds, summation order in the
GEMM/FFT and the order of atomic additions.  Integer and index outputs must be
bit-identical.
"""
import sys

import numpy as np
import pytest

from tests._cases import C, F, Grid, particles, rel_l2, rel_max, rng

pytestmark = pytest.mark.gpu

TOL_ELEM = 1e-14     # elementwise FP64 (FMA contraction only)
TOL_SUM = 1e-13      # sums over particles (atomic order)
TOL_FB = 1e-12       # DHT GEMM + FFT


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_library_loads(gpu):
    from chimera_amd._lib import lib
    assert b"gfx950" in lib().chimera_version()


def test_push_velocs_coords(gpu, port):
    say("case: push_velocs/push_coords np=5000")
    r = rng(100)
    u, f = F(r, 3, 5000), F(r, 6, 5000)
    a = gpu.push_velocs(u.copy(order="F"), f, 0.31)
    b = port.push_velocs(u.copy(order="F"), f, 0.31)
    assert rel_max(a, b) < TOL_ELEM
    x, xc = F(r, 3, 5000), F(r, 3, 5000)
    a = gpu.push_coords(x.copy(order="F"), u, xc.copy(order="F"), 0.05)
    b = port.push_coords(x.copy(order="F"), u, xc.copy(order="F"), 0.05)
    assert rel_max(a[0], b[0]) < TOL_ELEM and rel_max(a[1], b[1]) < TOL_ELEM


def test_push_empty(gpu):
    say("case: push_velocs np=0")
    u = np.zeros((3, 0), order="F")
    assert gpu.push_velocs(u, np.zeros((6, 0), order="F"), 0.1).shape == (3, 0)


@pytest.mark.parametrize("nko", [0, 1, 2, 3])
def test_deposition(gpu, port, nko):
    say(f"case: dep_curr/dep_dens nko={nko}")
    g = Grid(nx=64, nkr=16, nko=nko)
    coord, mom, w = particles(g, 6000, seed=200 + nko)
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr)
    z = np.zeros((g.nx, g.nkr + 1, nko + 1, 3), complex, order="F")
    a = gpu.dep_curr(coord, mom, w, z.copy(order="F"), *args)
    b = port.dep_curr(coord, mom, w, z.copy(order="F"), *args)
    assert rel_l2(a, b) < TOL_SUM
    assert not np.any(a[:, 0])
    zs = np.zeros((g.nx, g.nkr + 1, nko + 1), complex, order="F")
    a = gpu.dep_dens(coord, w, zs.copy(order="F"), *args)
    b = port.dep_dens(coord, w, zs.copy(order="F"), *args)
    assert rel_l2(a, b) < TOL_SUM


def test_deposition_chunked(gpu, port):
    say("case: dep_curr_chnk / dep_dens_chnk")
    g = Grid(nx=64, nkr=16, nko=1)
    coord, mom, w = particles(g, 3000, seed=210)
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr)
    ind = np.array([100, 900, 2000, 2900], dtype=np.int32)
    z = np.zeros((g.nx, g.nkr + 1, 2, 3), complex, order="F")
    a = gpu.dep_curr_chnk(coord, mom, w, z.copy(order="F"), ind, 2, *args)
    sub = slice(100, 2900)
    b = port.dep_curr(np.asfortranarray(coord[:, sub]), np.asfortranarray(mom[:, sub]), w[sub].copy(),
                      z.copy(order="F"), *args)
    assert rel_l2(a, b) < TOL_SUM
    zs = np.zeros((g.nx, g.nkr + 1, 2), complex, order="F")
    a = gpu.dep_dens_chnk(coord, w, zs.copy(order="F"), ind, 2, *args)
    b = port.dep_dens(np.asfortranarray(coord[:, sub]), w[sub].copy(), zs.copy(order="F"), *args)
    assert rel_l2(a, b) < TOL_SUM


def test_out_of_grid_particles_are_skipped(gpu):
    """A particle whose cell lies outside the arrays is skipped (the reference
    would write outside curr); NaN positions likewise.  No fault, no deposit."""
    say("case: out-of-grid particles")
    g = Grid(nx=32, nkr=8, nko=1)
    coord = np.asfortranarray(np.array([[g.leftX - 1.0, g.leftX + 100.0, np.nan],
                                        [0.1, 0.1, 0.1], [0.1, 0.1, 0.1]]))
    mom = np.ones((3, 3), order="F")
    w = np.ones(3)
    z = np.zeros((g.nx, g.nkr + 1, 2, 3), complex, order="F")
    a = gpu.dep_curr(coord, mom, w, z, g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr)
    assert not np.any(a)
    ft = gpu.proj_fld(coord, w, np.ones((g.nx, g.nkr + 1, 2, 6), complex, order="F"), np.zeros((6, 3), order="F"),
                      g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr)
    assert not np.any(ft)


@pytest.mark.parametrize("nko", [0, 2])
def test_gather_and_eb_correction(gpu, port, nko):
    say(f"case: proj_fld/eb_correction nko={nko}")
    g = Grid(nx=64, nkr=16, nko=nko)
    coord, mom, w = particles(g, 5000, seed=300 + nko)
    EB = C(rng(301), g.nx, g.nkr + 1, nko + 1, 6)
    ft = F(rng(302), 6, coord.shape[1])
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr)
    a = gpu.proj_fld(coord, w, EB, ft.copy(order="F"), *args)
    b = port.proj_fld(coord, w, EB, ft.copy(order="F"), *args)
    assert rel_max(a, b) < 1e-13
    assert rel_max(gpu.eb_correction(EB.copy(order="F")), port.eb_correction(EB.copy(order="F"))) < TOL_ELEM


def test_undulator(gpu, port):
    say("case: undul_analytic")
    g = Grid(nx=64, nkr=16, nko=1)
    coord, _, _ = particles(g, 4000, seed=400)
    p = np.array([1.7, 0.9, -2.0, 3.5])
    ft = F(rng(401), 6, coord.shape[1])
    a = gpu.undul_analytic(coord, ft.copy(order="F"), 0.0, p)
    b = port.undul_analytic(coord, ft.copy(order="F"), 0.0, p)
    assert rel_max(a, b) < 1e-13
    say("case: undul_analytic_taper")
    pt = np.array([1.7, 0.9, -2.0, 3.5, 0.3])
    a = gpu.undul_analytic_taper(coord, ft.copy(order="F"), 0.0, pt)
    b = port.undul_analytic_taper(coord, ft.copy(order="F"), 0.0, pt)
    assert rel_max(a, b) < 1e-13
    for name, params in (("undul_mapped", [0.9, -1.2, 0.1]), ("undul_mapped_tap", [0.9, -1.2, 0.1, 3.1, 0.25])):
        say(f"case: {name}, 40-node map, particles over and around it")
        r = rng(402)
        x = r.uniform(-1.5, 3.1, 4000)
        c = np.asfortranarray(np.vstack([x, r.normal(0, 0.3, x.size), r.normal(0, 0.3, x.size)]))
        a0, f0 = F(r, 2, 40), F(r, 6, x.size)
        a = getattr(gpu, name)(c, f0.copy(order="F"), 0.0, a0, np.array(params))
        b = getattr(port, name)(c, f0.copy(order="F"), 0.0, a0, np.array(params))
        assert rel_max(a, b) < 1e-13
    say("case: planewave, gaussbeam")
    pw = np.array([0.8, 1.3, 0.5, 7.0, 1.1, 0.35, 0.2])
    gb = np.array([1.3, 0.9, 3.0, 0.2, -0.1, 1.5, 0.8, 1.1])
    assert rel_max(gpu.planewave(coord, ft.copy(order="F"), 0.7, pw), port.planewave(coord, ft.copy(order="F"), 0.7, pw)) < 1e-13
    assert rel_max(gpu.gaussbeam(coord, ft.copy(order="F"), 0.4, 1.7, gb),
                   port.gaussbeam(coord, ft.copy(order="F"), 0.4, 1.7, gb)) < 1e-13


def test_genparts(gpu, port):
    say("case: genparts")
    nx, nr = 40, 13
    Xg = -2.0 + 0.1 * np.arange(nx)
    Rg = 0.3 * (np.arange(nr) - 0.5)
    RP = rng(500).random((nx, nr))
    pX, pR, pO = np.mgrid[1:2:2j, 1:2:2j, 1:4:4j]
    PX, PR = (pX.ravel() - 0.5) / 2, (pR.ravel() - 0.5) / 2
    PO = np.exp(2j * np.pi * (pO.ravel() - 1) / 4)
    co = np.zeros((4, nx * nr * 16), order="F")
    a = gpu.genparts(co.copy(order="F"), Xg, Rg, RP, PX, PR, PO)
    b = port.genparts(co.copy(order="F"), Xg, Rg, RP, PX, PR, PO)
    assert a[1] == b[1]
    assert np.abs(a[0] - b[0]).max() < 1e-14 * np.abs(b[0]).max()


def test_particle_bookkeeping(gpu, port):
    say("case: sortpartsout/chunk_coords_boundaries/align_data")
    g = Grid(nx=64, nkr=16, nko=1)
    coord, mom, w = particles(g, 7000, seed=600)
    lims = np.array([g.leftX + 0.5, g.leftX + 5.0, 0.0, 9.0])
    a, b = gpu.sortpartsout(coord, lims), port.sortpartsout(coord, lims)
    assert a[1] == b[1] and np.array_equal(a[0], b[0])
    Xg0 = g.leftX + g.dx * np.arange(g.nx + 1)
    a, b = gpu.chunk_coords_boundaries(coord, lims, Xg0, 4), port.chunk_coords_boundaries(coord, lims, Xg0, 4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    idx = np.argsort(a[0], kind="stable")[a[2]:]
    assert np.array_equal(gpu.align_data_vec(coord.copy(order="F"), idx), port.align_data_vec(coord.copy(order="F"), idx))
    assert np.array_equal(gpu.align_data_scl(w.copy(), idx), port.align_data_scl(w.copy(), idx))
    with pytest.raises(RuntimeError):
        gpu.align_data_scl(w.copy(), np.array([0, 1, 10 ** 7]))


# 2048x128: 32- and 64-column GEMM tiles of 64-complex x, and 2048-point
# complex128 lines whose twiddles stay in memory; 1024x64: twiddles staged in
# LDS and the phase formed by its own launch; 512x64: both in LDS; 8192: one
# line per workgroup (EPT = 32)
@pytest.mark.parametrize("nkx,nr,nkr,M", [(64, 16, 16, 3), (512, 64, 64, 2), (8, 5, 7, 1), (4096, 12, 12, 1),
                                          (2048, 128, 128, 2), (1024, 64, 64, 3), (8192, 16, 16, 1)])
def test_fb_transforms(gpu, port, nkx, nr, nkr, M):
    say(f"case: fb_vec_in/fb_scl_in/fb_eb_out nkx={nkx} nr={nr} nkr={nkr} M={M}")
    r = rng(700 + nkx)
    vec = C(r, nkx, nr + 1, M, 3)
    In = F(r, nr, nkr, M)
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.05)
    z = np.zeros((nkx, nkr, M, 3), complex, order="F")
    a = gpu.fb_vec_in(z.copy(order="F"), vec, -12.3, kx, In)
    b = port.fb_vec_in(z.copy(order="F"), vec, -12.3, kx, In)
    assert rel_l2(a, b) < TOL_FB
    s = vec[..., 1].copy(order="F")
    a = gpu.fb_scl_in(z[..., 0].copy(order="F"), s, -12.3, kx, In)
    b = port.fb_scl_in(z[..., 0].copy(order="F"), s, -12.3, kx, In)
    assert rel_l2(a, b) < TOL_FB
    e, bb = C(r, nkx, nkr, M, 6), C(r, nkx, nkr, M, 3)
    Out = F(r, nkr, nr, M)
    eb = np.ones((nkx, nr + 1, M, 6), complex, order="F")
    a = gpu.fb_eb_out(eb.copy(order="F"), e, bb, 3.1, kx, Out)
    b = port.fb_eb_out(eb.copy(order="F"), e, bb, 3.1, kx, Out)
    assert rel_l2(a, b) < TOL_FB
    assert not np.any(a[:, 0])


def test_fb_filtr(gpu, port):
    say("case: fb_filtr")
    r = rng(800)
    nkx, nkr, M = 256, 8, 2
    v = C(r, nkx, nkr, M, 3)
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.05)
    N = 40
    i = np.arange(N)
    filt = (i >= 0.75 * N) * (0.5 - 0.5 * np.cos(np.pi * (i - 0.75 * N) / (0.25 * N))) ** 2
    a = gpu.fb_filtr(v.copy(order="F"), -7.7, kx, filt, 0)
    b = port.fb_filtr(v.copy(order="F"), -7.7, kx, filt, 0)
    assert rel_l2(a, b) < TOL_FB
    with pytest.raises(RuntimeError):
        gpu.fb_filtr(v.copy(order="F"), -7.7, kx, filt, 2)


@pytest.mark.parametrize("nko", [0, 1, 3])
@pytest.mark.parametrize("nkx,nkr", [(64, 16), (128, 70)])
def test_fb_operators(gpu, port, nko, nkx, nkr):
    say(f"case: fb_rot/grad/div/graddiv nko={nko} nkx={nkx} nkr={nkr}")
    r = rng(900 + nko + nkr)
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.1)
    Dp, Dm = F(r, nkr, nkr, nko + 2), F(r, nkr, nkr, nko + 2)
    v, s = C(r, nkx, nkr, nko + 1, 3), C(r, nkx, nkr, nko + 1)
    z3 = np.zeros((nkx, nkr, nko + 1, 3), complex, order="F")
    z1 = np.zeros((nkx, nkr, nko + 1), complex, order="F")
    for name, out, inp in (("fb_rot", z3, v), ("fb_grad", z3, s), ("fb_div", z1, v)):
        a = getattr(gpu, name)(out.copy(order="F"), inp, Dp, Dm, kx)
        b = getattr(port, name)(out.copy(order="F"), inp, Dp, Dm, kx)
        assert rel_l2(a, b) < TOL_FB, name
    a = gpu.fb_graddiv(v.copy(order="F"), Dp, Dm, kx)
    b = port.fb_graddiv(v.copy(order="F"), Dp, Dm, kx)
    assert rel_l2(a, b) < TOL_FB


def test_maxwell(gpu, port):
    say("case: maxwell_* / poiss_corr / omp_*")
    r = rng(1000)
    sh = (64, 16, 3)
    EG = C(r, *sh, 6)
    j, gn, gp, gd = C(r, *sh, 3), C(r, *sh, 3), C(r, *sh, 3), C(r, *sh, 3)
    C1, C2 = F(r, *sh, 5), F(r, *sh, 5)
    assert rel_max(gpu.maxwell_push_with_spchrg(EG.copy(order="F"), j, gn, gp, C1, C2),
                   port.maxwell_push_with_spchrg(EG.copy(order="F"), j, gn, gp, C1, C2)) < 1e-13
    a3, b3 = np.asfortranarray(C1[..., :3]), np.asfortranarray(C2[..., :3])
    assert rel_max(gpu.maxwell_push_wo_spchrg(EG.copy(order="F"), j, a3, b3),
                   port.maxwell_push_wo_spchrg(EG.copy(order="F"), j, a3, b3)) < 1e-13
    CC1, CC2 = C(r, *sh, 2), C(r, *sh, 2)
    assert rel_max(gpu.maxwell_init_push(EG.copy(order="F"), j, gn, CC1, CC2),
                   port.maxwell_init_push(EG.copy(order="F"), j, gn, CC1, CC2)) < 1e-13
    w2 = F(r, *sh)
    assert rel_max(gpu.poiss_corr(j.copy(order="F"), gd, gn, gp, 20.0, w2),
                   port.poiss_corr(j.copy(order="F"), gd, gn, gp, 20.0, w2)) < 1e-13
    assert rel_max(gpu.omp_mult_vec(j.copy(order="F"), w2), port.omp_mult_vec(j.copy(order="F"), w2)) < TOL_ELEM
    assert rel_max(gpu.omp_add_vec(j.copy(order="F"), gn), port.omp_add_vec(j.copy(order="F"), gn)) < TOL_ELEM
    s = C(r, *sh)
    assert rel_max(gpu.omp_mult_scl(s.copy(order="F"), w2), port.omp_mult_scl(s.copy(order="F"), w2)) < TOL_ELEM
    assert rel_max(gpu.omp_add_scl(s.copy(order="F"), s), port.omp_add_scl(s.copy(order="F"), s)) < TOL_ELEM


def test_static_kick(gpu, port):
    """poiss_corr_stat and field_drift (the static-kick branch of the field
    update): drop-in FP64 to 1e-14 (FMA contraction; the device sin/cos of the
    drift phase are within an ulp of the host's), and the `_c8` entry points on
    device arrays to 1e-6 (float32 storage of fields and w2_inv; DT and kx stay
    FP64)."""
    from chimera_amd._lib import check, lib
    from chimera_amd.device import DeviceArray
    say("case: poiss_corr_stat / field_drift, 64x16x3, drop-in and _c8")
    r = rng(1010)
    sh = (64, 16, 3)
    EG, j, gn, gd = C(r, *sh, 6), C(r, *sh, 3), C(r, *sh, 3), C(r, *sh, 3)
    DT, w2 = C(r, sh[0]), F(r, *sh)
    kx = 2 * np.pi * np.fft.fftfreq(sh[0], 0.07)
    ref_j = port.poiss_corr_stat(j.copy(order="F"), gd, gn, DT, w2)
    ref_e = port.field_drift(EG.copy(order="F"), kx, 0.999, 0.035)
    assert rel_max(gpu.poiss_corr_stat(j.copy(order="F"), gd, gn, DT, w2), ref_j) < TOL_ELEM
    assert rel_max(gpu.field_drift(EG.copy(order="F"), kx, 0.999, 0.035), ref_e) < TOL_ELEM
    L = lib()
    dev = lambda a, t: DeviceArray.from_numpy(np.asfortranarray(a).astype(t))
    d_j, d_gd, d_gn, d_w2 = dev(j, np.complex64), dev(gd, np.complex64), dev(gn, np.complex64), dev(w2, np.float32)
    d_DT, d_EG, d_kx = dev(DT, np.complex128), dev(EG, np.complex64), dev(kx, np.float64)
    check("chimera_poiss_corr_stat_c8",
          L.chimera_poiss_corr_stat_c8(d_j.ptr, d_gd.ptr, d_gn.ptr, d_DT.ptr, d_w2.ptr, *sh))
    check("chimera_field_drift_c8", L.chimera_field_drift_c8(d_EG.ptr, d_kx.ptr, 0.999, 0.035, *sh))
    check("chimera_synchronize", L.chimera_synchronize())
    assert rel_max(d_j.numpy().astype(np.complex128), ref_j) < 1e-6
    assert rel_max(d_EG.numpy().astype(np.complex128), ref_e) < 1e-6


def test_particles_outside_the_grid(gpu, port):
    """Particles before leftX, on or past the last x node and at Rgrid[nr]
    are skipped by deposition and gather on both sides."""
    say("case: dep_curr/dep_dens/proj_fld with particles outside the grid")
    g = Grid(nx=40, nkr=10, nko=1)
    coord, mom, w = particles(g, 400, seed=64)
    xmax = g.leftX + (g.nx - 1) * g.dx
    out = np.array([[g.leftX - 0.3 * g.dx, 0.2, 0.1], [xmax, 0.2, 0.1], [xmax + 0.7 * g.dx, 0.1, 0.3],
                    [g.leftX + 5 * g.dx, 0.0, g.Rgrid[-1]]]).T
    coord = np.asfortranarray(np.concatenate([coord, out], axis=1))
    mom = np.asfortranarray(np.concatenate([mom, np.ones((3, 4))], axis=1))
    w = np.concatenate([w, np.ones(4)])
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr)
    z = np.zeros((g.nx, g.nkr + 1, 2, 3), complex, order="F")
    assert rel_l2(gpu.dep_curr(coord, mom, w, z.copy(order="F"), *args),
                  port.dep_curr(coord, mom, w, z.copy(order="F"), *args)) < TOL_SUM
    zr = np.zeros((g.nx, g.nkr + 1, 2), complex, order="F")
    assert rel_l2(gpu.dep_dens(coord, w, zr.copy(order="F"), *args),
                  port.dep_dens(coord, w, zr.copy(order="F"), *args)) < TOL_SUM
    Fld = C(rng(65), g.nx, g.nkr + 1, 2, 6)
    z6 = np.zeros((6, coord.shape[1]), order="F")
    a = gpu.proj_fld(coord, w, Fld, z6.copy(order="F"), *args)
    b = port.proj_fld(coord, w, Fld, z6.copy(order="F"), *args)
    assert np.all(a[:, -4:] == 0.0) and rel_max(a, b) < TOL_ELEM
