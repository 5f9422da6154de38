"""GPU parity of the envelope (KxShift) drop-in entry points of
grid_deps_env.f90, grid_deps_env_chnk.f90 and fb_math_env.f90, and of the
complex-table PSATD push of maxwell_solvers.f90:62-96, against the oracle
(bit-identical to the reference Fortran for these routines), through the C
ABI.

Mode axes -nkO..nkO with nkO = 0, 1, 2 and a carrier kx0 != 0.  Bars as for
the plain drop-ins: relative L2 <= 1e-13 for sums over particles (atomic
order), <= 1e-14 elementwise (FMA contraction only), <= 1e-12 for the
operators (the D+- products).  Particles on the axis and with zero weight or
zero momentum are among the inputs.
"""
import ctypes
import sys

import numpy as np
import pytest

from tests._cases import C, F, Grid, particles, rel_l2, rel_max, rng

pytestmark = pytest.mark.gpu

TOL_ELEM = 1e-14
TOL_SUM = 1e-13
TOL_OP = 1e-12
KX0 = 2.7


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("nko", [0, 1, 2])
def test_env_deposits(gpu, port, nko):
    say(f"case: dep_curr_env/dep_dens_env nko={nko}")
    g = Grid(nx=64, nkr=16, nko=nko)
    coord, mom, w = particles(g, 6000, seed=500 + nko)
    nm = 2 * nko + 1
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr, KX0)
    z = np.zeros((g.nx, g.nkr + 1, nm, 3), complex, order="F")
    a = gpu.dep_curr_env(coord, mom, w, z.copy(order="F"), *args)
    b = port.dep_curr_env(coord, mom, w, z.copy(order="F"), *args)
    say(f"  curr {rel_l2(a, b):.2e}")
    assert rel_l2(a, b) < TOL_SUM
    assert not np.any(a[..., :2]) and np.any(a[..., 2])
    zs = np.zeros((g.nx, g.nkr + 1, nm), complex, order="F")
    a = gpu.dep_dens_env(coord, w, zs.copy(order="F"), *args)
    b = port.dep_dens_env(coord, w, zs.copy(order="F"), *args)
    say(f"  dens {rel_l2(a, b):.2e}")
    assert rel_l2(a, b) < TOL_SUM


def test_env_deposits_chunked(gpu, port):
    """The chunked deposits of grid_deps_env_chnk.f90 cover particles
    IndInChunk(0)+1 .. IndInChunk(nchnk) (1-based): the plain envelope
    deposit of that slice."""
    say("case: dep_curr_env_chnk / dep_dens_env_chnk nko=1")
    g = Grid(nx=64, nkr=16, nko=1)
    coord, mom, w = particles(g, 3000, seed=520)
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr, KX0)
    ind = np.array([100, 900, 2000, 2900], dtype=np.int32)
    sub = slice(100, 2900)
    z = np.zeros((g.nx, g.nkr + 1, 3, 3), complex, order="F")
    a = gpu.dep_curr_env_chnk(coord, mom, w, z.copy(order="F"), ind, 2, *args)
    b = port.dep_curr_env(np.asfortranarray(coord[:, sub]), np.asfortranarray(mom[:, sub]), w[sub].copy(),
                          z.copy(order="F"), *args)
    assert rel_l2(a, b) < TOL_SUM
    zs = np.zeros((g.nx, g.nkr + 1, 3), complex, order="F")
    a = gpu.dep_dens_env_chnk(coord, w, zs.copy(order="F"), ind, 2, *args)
    b = port.dep_dens_env(np.asfortranarray(coord[:, sub]), w[sub].copy(), zs.copy(order="F"), *args)
    assert rel_l2(a, b) < TOL_SUM


@pytest.mark.parametrize("nko", [0, 2])
def test_env_gather_and_eb_correction(gpu, port, nko):
    say(f"case: proj_fld_env/eb_correction_env nko={nko}")
    g = Grid(nx=64, nkr=16, nko=nko)
    nm = 2 * nko + 1
    coord, mom, w = particles(g, 5000, seed=540 + nko)
    EB = C(rng(541), g.nx, g.nkr + 1, nm, 6)
    ft = F(rng(542), 6, coord.shape[1])
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr, KX0)
    a = gpu.proj_fld_env(coord, w, EB, ft.copy(order="F"), *args)
    b = port.proj_fld_env(coord, w, EB, ft.copy(order="F"), *args)
    say(f"  gather {rel_max(a, b):.2e}")
    assert rel_max(a, b) < TOL_ELEM
    a = gpu.eb_correction_env(EB.copy(order="F"))
    b = port.eb_correction_env(EB.copy(order="F"))
    assert np.array_equal(a, b)


def test_env_out_of_grid_particles_are_skipped(gpu):
    say("case: envelope deposits, out-of-grid particles")
    g = Grid(nx=32, nkr=8, nko=1)
    coord = np.asfortranarray(np.array([[g.leftX - 1.0, g.leftX + 100.0, np.nan],
                                        [0.1, 0.1, 0.1], [0.1, 0.1, 0.1]]))
    mom, w = np.ones((3, 3), order="F"), np.ones(3)
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr, KX0)
    assert not np.any(gpu.dep_curr_env(coord, mom, w, np.zeros((g.nx, g.nkr + 1, 3, 3), complex, order="F"), *args))
    assert not np.any(gpu.dep_dens_env(coord, w, np.zeros((g.nx, g.nkr + 1, 3), complex, order="F"), *args))
    ft = gpu.proj_fld_env(coord, w, np.ones((g.nx, g.nkr + 1, 3, 6), complex, order="F"), np.zeros((6, 3), order="F"),
                          *args)
    assert not np.any(ft)


@pytest.mark.parametrize("nko", [0, 1, 2])
def test_env_operators(gpu, port, nko):
    say(f"case: fb_rot_env/fb_grad_env/fb_div_env/fb_graddiv_env nko={nko}")
    r = rng(560 + nko)
    nkx, nkr = 96, 24
    nm = 2 * nko + 1
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.1) + KX0
    Dp, Dm = F(r, nkr, nkr, nm + 2), F(r, nkr, nkr, nm + 2)
    v, s = C(r, nkx, nkr, nm, 3), C(r, nkx, nkr, nm)
    z3 = np.zeros((nkx, nkr, nm, 3), complex, order="F")
    z1 = np.zeros((nkx, nkr, nm), complex, order="F")
    for name, out, inp in (("fb_rot_env", z3, v), ("fb_grad_env", z3, s), ("fb_div_env", z1, v)):
        a = getattr(gpu, name)(out.copy(order="F"), inp, Dp, Dm, kx)
        b = getattr(port, name)(out.copy(order="F"), inp, Dp, Dm, kx)
        say(f"  {name} {rel_l2(a, b):.2e}")
        assert rel_l2(a, b) < TOL_OP, name
    a = gpu.fb_graddiv_env(v.copy(order="F"), Dp, Dm, kx)
    b = port.fb_graddiv_env(v.copy(order="F"), Dp, Dm, kx)
    say(f"  fb_graddiv_env {rel_l2(a, b):.2e}")
    assert rel_l2(a, b) < TOL_OP


def test_env_operators_refuse_even_mode_axis(gpu):
    """The envelope mode axis is -nkO..nkO: an even length is refused before
    any launch."""
    nkx, nkr = 8, 4
    with pytest.raises(ValueError):
        gpu.fb_div_env(np.zeros((nkx, nkr, 2), complex, order="F"), np.zeros((nkx, nkr, 2, 3), complex, order="F"),
                       np.zeros((nkr, nkr, 4), order="F"), np.zeros((nkr, nkr, 4), order="F"), np.zeros(nkx))


def test_env_psatd_complex_tables(gpu, port):
    say("case: maxwell_push_wo_spchrg, complex tables")
    r = rng(580)
    sh = (64, 16, 5)
    EG, j = C(r, *sh, 6), C(r, *sh, 3)
    C1, C2 = C(r, *sh, 3), C(r, *sh, 3)
    a = gpu.maxwell_push_wo_spchrg(EG.copy(order="F"), j, C1, C2)
    b = port.maxwell_push_wo_spchrg(EG.copy(order="F"), j, C1, C2)
    say(f"  {rel_max(a, b):.2e}")
    assert rel_max(a, b) < TOL_ELEM


@pytest.mark.parametrize("nko", [0, 1])
def test_env_resident_entries(gpu, port, nko):
    """The resident envelope entries (particles structure-of-arrays in device
    memory) against the oracle: the deposits are left unfolded (the forward
    transform folds them), so the ghost fold is applied here on the host; the
    gather runs with an analytic undulator and the momentum push fused, the
    oracle as proj_fld_env, undul_analytic, push_velocs."""
    from chimera_amd.device import DeviceArray, call
    from chimera_amd.species import DeviceRecord
    say(f"case: soa_dep_curr_env / soa_dep_dens_env / soa_gather_push_env nko={nko}")
    g = Grid(nx=64, nkr=16, nko=nko)
    coord, mom, w = particles(g, 5000, seed=590 + nko)
    n, nm = coord.shape[1], 2 * nko + 1
    X = DeviceArray.from_numpy(np.asfortranarray(coord.T))
    U = DeviceArray.from_numpy(np.asfortranarray(mom.T))
    Wd = DeviceArray.from_numpy(w)
    R = DeviceArray.from_numpy(g.Rgrid)
    geom = (g.leftX, R.ptr, 1 / g.dx, 1 / g.dr, float(g.Rgrid[0]), float(g.Rgrid[-1]), KX0, g.nx - 1, g.nkr, nko)
    args = (g.leftX, g.Rgrid, 1 / g.dx, 1 / g.dr, KX0)

    def fold(a):
        a[:, 1] -= a[:, 0]
        a[:, 0] = 0.0
        return a

    J = DeviceArray((g.nx, g.nkr + 1, nm, 3), np.complex128)
    call("chimera_soa_dep_curr_env", X.ptr, U.ptr, Wd.ptr, n, n, J.ptr, *geom, None)
    a = fold(J.numpy())
    b = port.dep_curr_env(coord, mom, w, np.zeros((g.nx, g.nkr + 1, nm, 3), complex, order="F"), *args)
    say(f"  curr {rel_l2(a, b):.2e}")
    assert rel_l2(a, b) < TOL_SUM
    Dn = DeviceArray((g.nx, g.nkr + 1, nm), np.complex128)
    call("chimera_soa_dep_dens_env", X.ptr, Wd.ptr, n, n, Dn.ptr, *geom, None)
    a = fold(Dn.numpy())
    b = port.dep_dens_env(coord, w, np.zeros((g.nx, g.nkr + 1, nm), complex, order="F"), *args)
    say(f"  dens {rel_l2(a, b):.2e}")
    assert rel_l2(a, b) < TOL_SUM

    EB = C(rng(592), g.nx, g.nkr + 1, nm, 6)
    EBd = DeviceArray.from_numpy(EB)
    und = np.array([1.3, 2.0, g.leftX - 1.0, 40.0])
    recs = (DeviceRecord * 1)()
    recs[0].kind = 0
    for i, v in enumerate(und):
        recs[0].params[i] = v
    push_dt, t = 0.37, 1.5
    call("chimera_soa_gather_push_env", X.ptr, U.ptr, Wd.ptr, n, n, EBd.ptr, *geom, push_dt, ctypes.addressof(recs), 1,
         t, None)
    a = U.numpy().T
    ft = port.proj_fld_env(coord, w, EB, np.zeros((6, n), order="F"), *args)
    ft = port.undul_analytic(coord, ft, t, und)
    b = port.push_velocs(mom.copy(order="F"), ft, push_dt)
    say(f"  gather+push {rel_max(a, b):.2e}")
    assert rel_max(a, b) < TOL_SUM
    assert rel_l2(a, mom) > 1e-3                      # the push acted
