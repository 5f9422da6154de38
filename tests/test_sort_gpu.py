"""The counting sort of particles by cell (chimera_soa_sort_cells).

Bit-exact bar: the permutation equals numpy.argsort(keys, kind='stable') of the
cell keys recomputed on the host from the grid_deps.f90:46-47 expressions,
particles outside the domain are dropped, particles outside the grid are kept
after all cells, the gathered arrays are exact copies, and two runs give
identical bytes.  Then a whole run with sorting every step must give the same
fields as one without sorting, up to the order of floating-point sums.
"""
import ctypes
import sys

import numpy as np
import pytest

from tests._cases import Grid, particles, rel_l2

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def host_keys(coord, g, lims):
    x, y, z = coord
    nx, nr = g.nx - 1, g.nkr
    rp = np.sqrt(y * y + z * z)
    fx = (x - g.leftX) * (1 / g.dx)
    fr = (rp - g.Rgrid[0]) * (1 / g.dr)
    inside = (fx >= 0) & (fx < nx) & (fr >= 0) & (fr < nr)
    key = np.where(inside, np.floor(fr) * nx + np.floor(fx), nx * nr).astype(np.int64)
    if lims is not None:
        r2 = y * y + z * z
        keep = (x >= lims[0]) & (x <= lims[1]) & (r2 >= lims[2]) & (r2 <= lims[3])
        key = np.where(keep, key, nx * nr + 1)
    return key, nx * nr


def _sort_once(coord, mom, w, g, lims):
    from chimera_amd.device import DeviceArray, call
    n = coord.shape[1]
    cap = n + 100
    src = {k: DeviceArray((cap, 3), np.float64) for k in ("x", "xh", "u")}
    dst = {k: DeviceArray((cap, 3), np.float64) for k in ("x", "xh", "u")}
    sw, dw = DeviceArray((cap,), np.float64), DeviceArray((cap,), np.float64)
    for c in range(3):
        src["x"].write(c * cap, coord[c])
        src["xh"].write(c * cap, coord[c] + 0.25)
        src["u"].write(c * cap, mom[c])
    sw.write(0, w)
    perm = DeviceArray((n,), np.int32)
    nout = ctypes.c_int32(0)
    L = np.ascontiguousarray(lims, dtype=np.float64) if lims is not None else None
    call("chimera_soa_sort_cells", src["x"].ptr, src["xh"].ptr, src["u"].ptr, sw.ptr, cap, n, dst["x"].ptr,
         dst["xh"].ptr, dst["u"].ptr, dw.ptr, cap, float(g.leftX), 1 / g.dx, 1 / g.dr, float(g.Rgrid[0]), g.nx - 1,
         g.nkr, L.ctypes.data if L is not None else None, 0, ctypes.addressof(nout), perm.ptr)
    k = nout.value
    out = {key: dst[key].numpy()[:k].T for key in dst}
    out["w"] = dw.numpy()[:k]
    out["perm"] = perm.numpy()
    out["k"] = k
    return out


@pytest.mark.parametrize("with_lims", [False, True])
def test_sort_is_stable_argsort(with_lims):
    say(f"case: sort_cells with_lims={with_lims}")
    g = Grid(nx=96, nkr=24, nko=1)
    coord, mom, w = particles(g, 20000, seed=77)
    # shuffle, duplicate a block (equal keys), and put some particles outside the grid
    rng = np.random.default_rng(3)
    coord = np.asfortranarray(coord[:, rng.permutation(coord.shape[1])])
    coord[:, 100:200] = coord[:, 300:400]
    coord[0, 500:510] = g.leftX - 5.0
    coord[1, 600:610] = 1e3
    lims = np.array([g.leftX + 1.0, g.leftX + 7.0, 0.0, 25.0]) if with_lims else None
    key, ncell = host_keys(coord, g, lims)
    order = np.argsort(key, kind="stable")
    kept = order[key[order] <= ncell]
    a = _sort_once(coord, mom, w, g, lims)
    assert a["k"] == kept.size
    assert np.array_equal(a["perm"][:a["k"]], kept)
    assert np.array_equal(a["x"], coord[:, kept]) and np.array_equal(a["u"], mom[:, kept])
    assert np.array_equal(a["xh"], coord[:, kept] + 0.25) and np.array_equal(a["w"], w[kept])
    b = _sort_once(coord, mom, w, g, lims)
    assert np.array_equal(a["perm"], b["perm"]) and np.array_equal(a["x"], b["x"])


def test_sort_empty_and_single():
    say("case: sort_cells tiny inputs")
    g = Grid(nx=32, nkr=8, nko=0)
    coord, mom, w = particles(g, 1, seed=1)
    a = _sort_once(coord, mom, w, g, None)
    assert a["k"] == 1 and np.array_equal(a["x"], coord)


def test_sorting_keeps_the_physics():
    """A run that sorts every step against one that never sorts."""
    import chimera_amd as G
    from chimera_amd import workloads as W
    say("case: lpa_small with SortEvery=1 vs no sorting, 8 steps")
    spec = W.lpa_small(Nx=128, Nkr=32, nko=1)
    a = W.build(G, spec, 3)
    b = W.build(G, spec, 3, SortEvery=1)
    for i in range(1, 9):
        a.make_step(i)
        b.make_step(i)
    for key in ("EG_fb", "EB"):
        err = rel_l2(b.Solvers[0].Data[key].numpy(), a.Solvers[0].Data[key].numpy())
        say(f"  {key}: {err:.2e}")
        assert err < 1e-11
    ua, ub = a.Particles[0].host("momenta"), b.Particles[0].host("momenta")
    assert ua.shape == ub.shape
    # same particles, other order: compare sorted momenta
    assert rel_l2(np.sort(ub.ravel()), np.sort(ua.ravel())) < 1e-11


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sorting_on_window_moves_keeps_the_physics(precision):
    """A moving-window run (injection and a 16-cell absorbing layer every 10
    steps) that sorts by cell on every move, where the sort also drops the
    particles in the layer in place of the compaction, against the same run
    that never sorts: the same particle count, fields and sorted momenta
    (FP64: 1e-11; mixed: 1e-5)."""
    import chimera_amd as G
    from chimera_amd import workloads as W
    say(f"case: lpa_small moving window, SortEvery=10 vs no sorting, 31 steps, {precision}")
    spec = W.with_precision(W.lpa_small(Nx=256, Nkr=32, nko=1, moving=True, steps=31), precision)
    spec["frames"][0]["AbsorbLayer"] = 16
    # each run draws its injected plasma from its own stream, seeded alike
    a = W.build(G, spec, 3, own_rng=True)
    b = W.build(G, spec, 3, own_rng=True, SortEvery=10)
    for i in range(1, 32):
        a.make_step(i)
        b.make_step(i)
        assert a.Particles[0].np == b.Particles[0].np, i
    tol = 1e-11 if precision == "fp64" else 1e-5
    for key in ("EG_fb", "EB"):
        err = rel_l2(b.Solvers[0].Data[key].numpy().astype(np.complex128),
                     a.Solvers[0].Data[key].numpy().astype(np.complex128))
        say(f"  {key}: {err:.2e}")
        assert err < tol
    ua = a.Particles[0].host("momenta").astype(np.float64)
    ub = b.Particles[0].host("momenta").astype(np.float64)
    assert rel_l2(np.sort(ub.ravel()), np.sort(ua.ravel())) < tol
