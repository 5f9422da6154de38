"""The GEMM tile autotuner (chimera_set_autotune): with it on, the first
launch of each GEMM shape times the tile shapes and keeps the fastest; the
results are the same bits as with the block-count rule (every tile shape sums
each output over K in the same order), a tuned shape is not tuned again, and
the tuned shapes written to the cache file are read back."""
import sys

import numpy as np
import pytest

from tests._cases import C, F, rng

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bits(a):
    """The array's bit pattern (complex128 -> uint64 pairs, complex64 -> uint32 pairs)."""
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.complex128 else np.uint32)


def _c8_op(op, out_shape, inp, Dp, Dm, kx, nko):
    from chimera_amd.device import DeviceArray, call
    d_in = DeviceArray.from_numpy(inp, np.complex64)
    d_out = DeviceArray(out_shape, np.complex64)
    d_p, d_m = DeviceArray.from_numpy(Dp, np.float32), DeviceArray.from_numpy(Dm, np.float32)
    d_kx = DeviceArray.from_numpy(kx)
    call("chimera_fb_op_c8", op, d_out.ptr, d_in.ptr, d_p.ptr, d_m.ptr, d_kx.ptr, inp.shape[0], inp.shape[1], nko)
    return d_out.numpy()


def test_autotuned_tiles_same_bits(gpu, tmp_path):
    from chimera_amd._lib import lib
    L = lib()
    r = rng(4242)
    nkx, nkr, nko = 2048, 96, 2
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.1)
    Dp, Dm = F(r, nkr, nkr, nko + 2), F(r, nkr, nkr, nko + 2)
    v = C(r, nkx, nkr, nko + 1, 3)
    z3 = np.zeros((nkx, nkr, nko + 1, 3), complex, order="F")
    path = str(tmp_path / "tiles.txt").encode()
    old_form = L.chimera_get_mixed_gemm()
    try:
        L.chimera_set_autotune(0, None)
        ref64 = gpu.fb_rot(z3.copy(order="F"), v, Dp, Dm, kx)
        ref8 = {}
        for form in (0, 1):
            L.chimera_set_mixed_gemm(form)
            ref8[form] = _c8_op(0, z3.shape, v, Dp, Dm, kx, nko)
        n0 = L.chimera_autotune_entries()
        assert L.chimera_set_autotune(1, path) == 0
        say("case: fb_rot 2048x96 m=0..2, autotuned: fp64, split-bf16, fp32")
        a64 = gpu.fb_rot(z3.copy(order="F"), v, Dp, Dm, kx)
        n1 = L.chimera_autotune_entries()
        assert n1 > n0
        assert np.array_equal(_bits(a64), _bits(ref64))
        for form in (0, 1):
            L.chimera_set_mixed_gemm(form)
            a8 = _c8_op(0, z3.shape, v, Dp, Dm, kx, nko)
            assert np.array_equal(_bits(a8), _bits(ref8[form])), form
        n2 = L.chimera_autotune_entries()
        gpu.fb_rot(z3.copy(order="F"), v, Dp, Dm, kx)
        assert L.chimera_autotune_entries() == n2               # tuned once per shape
        lines = open(path).read().splitlines()
        say(f"  {n2 - n0} shapes tuned; file lines: {len(lines)}")
        assert len(lines) == n2 - n0
        for ln in lines:
            f = ln.split()
            assert len(f) == 11 and 0 <= int(f[9]) <= 4 and float(f[10]) > 0
        assert L.chimera_set_autotune(1, path) == 0            # re-reads the file: nothing new to tune
        gpu.fb_rot(z3.copy(order="F"), v, Dp, Dm, kx)
        assert len(open(path).read().splitlines()) == len(lines)
    finally:
        L.chimera_set_autotune(0, None)
        L.chimera_set_mixed_gemm(old_form)
