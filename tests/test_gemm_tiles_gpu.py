"""The FB operators on grids large enough for every tile shape of the GEMM,
against the oracle, in FP64 (drop-in entry points) and mixed precision (the
resident `_c8` entry points on device arrays).

The small-grid parity tests (drop-in operators at 64x16 and 128x70, the
lpa_cfg0 steps, the bunches) run the 32-column tiles of 32- and 16-complex x;
the grids here make the launches choose the larger ones as well (modes 0..2):
1024x128 (32- and 16-complex x), 4096x128 (32- and 64-column tiles of
64-complex x, buffer-load stages) and 8192x256 (128-column tiles, with two
accumulators in FP32).  Operands are seeded random fields and matrices of the
reference's shapes (DpS2S/DmS2S of fb_math.f90: nkr x nkr x nko+2).

Bars: 1e-12 relative L2 in FP64 (summation order of the GEMM only; 2-5e-16
measured); 2e-6 in mixed precision, where the FP32 MFMA sums K = 128 or 256
products of float32-rounded operands (2-4e-7 measured).  The operands are
scaled so that the GEMM terms dominate the outputs.
"""
import sys

import numpy as np
import pytest

from tests._cases import C, F, rel_l2, rng

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6

CASES = [(1024, 128, 2, ("rot", "grad", "div", "graddiv")),
         (4096, 128, 2, ("rot", "grad", "div", "graddiv")),
         (8192, 256, 2, ("grad", "div"))]
OPS = {"rot": 0, "grad": 1, "div": 2, "graddiv": 3}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _c8(op, out, inp, Dp, Dm, kx, nko):
    """chimera_fb_op_c8 on device copies of the operands; returns the output."""
    from chimera_amd._lib import check, lib
    from chimera_amd.device import DeviceArray
    L = lib()
    d_out = DeviceArray.from_numpy(out.astype(np.complex64))
    d_in = d_out if op == OPS["graddiv"] else DeviceArray.from_numpy(inp.astype(np.complex64))
    d_p, d_m = DeviceArray.from_numpy(Dp.astype(np.float32)), DeviceArray.from_numpy(Dm.astype(np.float32))
    d_kx = DeviceArray.from_numpy(kx)
    nkx, nkr = inp.shape[0], inp.shape[1]
    check("chimera_fb_op_c8", L.chimera_fb_op_c8(op, d_out.ptr, d_in.ptr, d_p.ptr, d_m.ptr, d_kx.ptr, nkx, nkr, nko))
    check("chimera_synchronize", L.chimera_synchronize())
    return d_out.numpy().astype(np.complex128)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("nkx,nkr,nko,ops", CASES)
def test_operators_every_tile(gpu, port, nkx, nkr, nko, ops):
    r = rng(7000 + nkx + nkr + nko)
    # |kx| <= pi and unit-variance matrices: the GEMM terms (about sqrt(nkr) per
    # element) outweigh the diagonal i*kx terms, so the bars measure the GEMM
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 1.0)
    Dp, Dm = F(r, nkr, nkr, nko + 2), F(r, nkr, nkr, nko + 2)
    v, s = C(r, nkx, nkr, nko + 1, 3), C(r, nkx, nkr, nko + 1)
    z3 = np.zeros((nkx, nkr, nko + 1, 3), complex, order="F")
    z1 = np.zeros((nkx, nkr, nko + 1), complex, order="F")
    for name in ops:
        say(f"case: fb_{name} nkx={nkx} nkr={nkr} nko={nko}")
        if name == "graddiv":
            ref = port.fb_graddiv(v.copy(order="F"), Dp, Dm, kx)
            a64 = gpu.fb_graddiv(v.copy(order="F"), Dp, Dm, kx)
            a32 = _c8(OPS[name], v.copy(order="F"), v, Dp, Dm, kx, nko)
        else:
            out, inp = (z1, v) if name == "div" else (z3, s if name == "grad" else v)
            ref = getattr(port, "fb_" + name)(out.copy(order="F"), inp, Dp, Dm, kx)
            a64 = getattr(gpu, "fb_" + name)(out.copy(order="F"), inp, Dp, Dm, kx)
            a32 = _c8(OPS[name], out.copy(order="F"), inp, Dp, Dm, kx, nko)
        e64, e32 = rel_l2(a64, ref), rel_l2(a32, ref)
        say(f"  fp64 {e64:.2e}  mixed {e32:.2e}")
        assert e64 < TOL64, (name, e64)
        assert e32 < TOL32, (name, e32)


def _tile(nkx, maxN, njobs, has_w4, min_blocks=512):
    """The launch's tile as the library chooses it (gemm_tile in fb.hip), to
    show which shapes the cases reach; index 0..4 = 128-, 64-, 32-column tiles
    of 64-complex x, 32-column tiles of 32- and 16-complex x."""
    xb, c32 = -(-nkx // 64), -(-maxN // 32)
    nb = [xb * -(-maxN // 128) * njobs, xb * -(-maxN // 64) * njobs, xb * c32 * njobs, -(-nkx // 32) * c32 * njobs]
    for t in range(0 if has_w4 else 1, 4):
        if nb[t] >= min_blocks:
            return t
    return 4


def test_cases_reach_every_tile():
    """The operator launches of CASES (grad: nko+1 two-accumulator jobs; div:
    nko+1 one-accumulator jobs; the div of graddiv: nko+2) choose all five
    shapes, and the 128-column shape with two accumulators in FP32."""
    seen64, seen32 = set(), set()
    for nkx, nkr, nko, ops in CASES:
        n = nko + 1
        if "grad" in ops:
            seen64.add(_tile(nkx, nkr, n, False))
            seen32.add(("two", _tile(nkx, nkr, n, True)))
        if "div" in ops:
            seen64.add(_tile(nkx, nkr, n, True))
        if "graddiv" in ops:
            seen64.add(_tile(nkx, nkr, n + 1, True))
    assert seen64 | {t for _, t in seen32} == {0, 1, 2, 3, 4}, (seen64, seen32)
    assert ("two", 0) in seen32
