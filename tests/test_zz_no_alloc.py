"""chimera_beam_moments_noalloc (the entry point that allocates nothing)
against chimera_beam_moments, through the C ABI: the same bits on two calls in
a row, the second with fewer particles, then no particles; refusal of a host
output pointer."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sums(name, x, u, w, ld, n):
    from chimera_amd.device import DeviceArray, call
    out = DeviceArray((13,), np.float64)
    call(name, x.ptr, u.ptr, w.ptr, ld, n, out.ptr)
    return out.numpy()


def test_noalloc_matches_allocating_entry():
    from chimera_amd.device import DeviceArray
    r = np.random.default_rng(77)
    n0 = 300_000
    # structure of arrays: component c of particle i at c*n0 + i, i.e. (n0, 3) in Fortran order
    x = DeviceArray.from_numpy((r.normal(size=(3, n0)) * [[2.0], [0.5], [0.5]]).T)
    u = DeviceArray.from_numpy((r.normal(size=(3, n0)) + [[40.0], [0.0], [0.0]]).T)
    w = DeviceArray.from_numpy(r.uniform(0.1, 1.0, n0))
    for n in (n0, 1000, 0):
        print(f"case: beam_moments np={n} (ld={n0})", flush=True)
        a = _sums("chimera_beam_moments", x, u, w, n0, n)
        b = _sums("chimera_beam_moments_noalloc", x, u, w, n0, n)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (n, a, b)
        if n:
            xs, us, ws = x.numpy()[:n].T, u.numpy()[:n].T, w.numpy()[:n]
            assert np.isclose(b[0], ws.sum(), rtol=1e-12)
            assert np.isclose(b[1], (xs[0] * ws).sum(), rtol=1e-10)
            assert np.isclose(b[9], (us[2] ** 2 * ws).sum(), rtol=1e-10)
        else:
            assert not np.any(b)


def test_noalloc_refuses_host_output():
    from chimera_amd._lib import lib
    from chimera_amd.device import DeviceArray
    x = DeviceArray((10, 3), np.float64)
    u = DeviceArray((10, 3), np.float64)
    w = DeviceArray((10,), np.float64)
    host = np.zeros(13)
    st = lib().chimera_beam_moments_noalloc(x.ptr, u.ptr, w.ptr, 10, 10, host.ctypes.data)
    assert st != 0 and b"device memory" in lib().chimera_last_error()
