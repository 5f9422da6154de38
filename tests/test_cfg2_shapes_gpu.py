"""The Fourier-Bessel transforms and operators at the precision-form size
(Nx = 8192, Nkr = 512, modes m = 0..3) with the real Bessel operators of the
Solver (In/Out and DpS2S/DmS2S of solvers.py:717-759, axes swapped as
solvers.py:152-157), against the FP64 oracle, error per mode.

FP64 (drop-in entry points, fb_io.f90 and fb_math.f90 layouts): relative L2
<= 1e-12 per mode.  Mixed precision (the resident complex64 entry points on
device arrays) in both GEMM forms, split-BF16 (six products) and FP32 MFMA:
<= 1e-5 per mode, m = 3 (the worst-conditioned inverse Hankel matrix)
included.  Inputs: seeded fields whose radial profile is a smooth beam plus
1e-3 white noise, so that the ill-conditioned high-order inverse matrices
see the spectrum a real run gives them.
"""
import sys

import numpy as np
import pytest

from tests._cases import rel_l2

pytestmark = pytest.mark.gpu

NX, NKR, NKO, DX, DR = 8192, 512, 3, 0.05, 0.3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ops():
    from chimera_amd import operators as O
    kx = 2 * np.pi * np.fft.fftfreq(NX, DX)
    nkr, Rg, lR = O.r_grid(NKR * DR, DR)
    e0 = O.spectral_operators(NX, nkr, NKO, lR, kx, Rg, ext=0)
    e1 = O.spectral_operators(NX, nkr, NKO, lR, kx, Rg, ext=1)
    sw = lambda a: np.asfortranarray(np.swapaxes(a, 0, 1))      # noqa: E731
    return {"kx": kx, "In": sw(e0["In"]), "Out": sw(e0["Out"]), "Dp": sw(e1["DpS2S"]), "Dm": sw(e1["DmS2S"]),
            "R": Rg}


def _field(r, prof, shape_tail):
    """(NX, len(prof), ...) complex: the profile along the second axis times a
    random x signal, plus 1e-3 white noise."""
    n = int(np.prod(shape_tail))
    xs = r.standard_normal((NX, 1, n)) + 1j * r.standard_normal((NX, 1, n))
    f = prof[None, :, None] * xs + 1e-3 * (r.standard_normal((NX, prof.shape[0], n)) +
                                            1j * r.standard_normal((NX, prof.shape[0], n)))
    return np.asfortranarray(f.reshape((NX, prof.shape[0]) + tuple(shape_tail), order="C"))


def _per_mode(a, b, axis=2):
    return [rel_l2(np.take(a, m, axis=axis), np.take(b, m, axis=axis)) for m in range(a.shape[axis])]


def _c8_forward(vec, ops, leftX):
    from chimera_amd.device import DeviceArray, call
    nkx, nr1, M = vec.shape[:3]
    d_in = DeviceArray.from_numpy(vec, np.complex64)
    d_out = DeviceArray((nkx, NKR, M), np.complex64)
    d_kx, d_In = DeviceArray.from_numpy(ops["kx"]), DeviceArray.from_numpy(ops["In"], np.float32)
    call("chimera_fb_forward_c8", d_out.ptr, d_in.ptr, leftX, d_kx.ptr, d_In.ptr, None, nkx, nr1 - 1, M, NKR, 1, 0,
         None)
    return d_out.numpy().astype(np.complex128)


def _c8_inverse(e, b, ops, leftX):
    from chimera_amd.device import DeviceArray, call
    nkx, nkr, M = e.shape[:3]
    nr = ops["Out"].shape[1]
    d_e, d_b = DeviceArray.from_numpy(e, np.complex64), DeviceArray.from_numpy(b, np.complex64)
    d_eb = DeviceArray((nkx, nr + 1, M, 6), np.complex64)
    d_kx, d_Out = DeviceArray.from_numpy(ops["kx"]), DeviceArray.from_numpy(ops["Out"], np.float32)
    call("chimera_fb_inverse_c8", d_eb.ptr, d_e.ptr, d_b.ptr, leftX, d_kx.ptr, d_Out.ptr, nkx, nr, M, nkr, 1, None)
    return d_eb.numpy().astype(np.complex128)


def _c8_op(op, out_shape, inp, ops):
    from chimera_amd.device import DeviceArray, call
    d_in = DeviceArray.from_numpy(inp, np.complex64)
    d_out = DeviceArray(out_shape, np.complex64)
    d_p, d_m = DeviceArray.from_numpy(ops["Dp"], np.float32), DeviceArray.from_numpy(ops["Dm"], np.float32)
    d_kx = DeviceArray.from_numpy(ops["kx"])
    call("chimera_fb_op_c8", op, d_out.ptr, d_in.ptr, d_p.ptr, d_m.ptr, d_kx.ptr, NX, NKR, NKO)
    return d_out.numpy().astype(np.complex128)


FORMS = {"bf16x6": 0, "fp32": 1}


def _with_form(form, fn):
    from chimera_amd._lib import lib
    L = lib()
    old = L.chimera_get_mixed_gemm()
    assert L.chimera_set_mixed_gemm(FORMS[form]) == 0
    try:
        return fn()
    finally:
        L.chimera_set_mixed_gemm(old)


@pytest.mark.timeout(900)
def test_cfg2_hankel_transforms(gpu, port, ops):
    r = np.random.default_rng(2026)
    M = NKO + 1
    leftX = -409.55
    s = _field(r, np.exp(-(np.maximum(ops["R"], 0) / 12.0) ** 2), (M,))      # a beam of radius 12
    say("case: fb_scl_in Nx=8192 Nkr=512 m=0..3")
    z = np.zeros((NX, NKR, M), complex, order="F")
    ref = port.fb_scl_in(z.copy(order="F"), s, leftX, ops["kx"], ops["In"])
    a64 = gpu.fb_scl_in(z.copy(order="F"), s, leftX, ops["kx"], ops["In"])
    e64 = _per_mode(a64, ref)
    say("  fp64 per mode: " + " ".join(f"{e:.2e}" for e in e64))
    assert max(e64) < 1e-12
    for form in FORMS:
        a8 = _with_form(form, lambda: _c8_forward(s, ops, leftX))
        e8 = _per_mode(a8, ref)
        say(f"  mixed {form} per mode: " + " ".join(f"{e:.2e}" for e in e8))
        assert max(e8) < 1e-5, (form, e8)
    say("case: fb_eb_out Nx=8192 Nkr=512 m=0..3")
    # E and B spectra decaying with the Bessel index (a field the inverse
    # transform meets in a run), rounded to complex64 first so that the mixed
    # error measures the transform, not the rounding of its input: the
    # inverse Hankel sums at high modes cancel (|Out . e| << |Out| |e|), which
    # turns the 6e-8 rounding of e alone into ~1e-4 at m = 3 (printed)
    e = _field(r, np.exp(-np.arange(NKR) / 48.0), (M, 6)).astype(np.complex64).astype(np.complex128)
    e = np.asfortranarray(e)
    b = np.asfortranarray(e[..., 3:])
    nr = ops["Out"].shape[1]
    eb = np.zeros((NX, nr + 1, M, 6), complex, order="F")
    ref2 = port.fb_eb_out(eb.copy(order="F"), e, b, leftX, ops["kx"], ops["Out"])
    a64 = gpu.fb_eb_out(eb.copy(order="F"), e, b, leftX, ops["kx"], ops["Out"])
    e64 = _per_mode(a64, ref2)
    say("  fp64 per mode: " + " ".join(f"{x:.2e}" for x in e64))
    assert max(e64) < 1e-12
    for form in FORMS:
        a8 = _with_form(form, lambda: _c8_inverse(e, b, ops, leftX))
        e8 = _per_mode(a8, ref2)
        say(f"  mixed {form} per mode: " + " ".join(f"{x:.2e}" for x in e8))
        assert max(e8) < 1e-5, (form, e8)
    # the same with e from the forward transform of the beam (ill-conditioned at m = 3): reported
    e2 = np.asfortranarray(np.concatenate([ref[..., None]] * 3 + [1j * ref[..., None]] * 3, axis=3))
    ref3 = port.fb_eb_out(eb.copy(order="F"), e2, np.asfortranarray(e2[..., 3:]), leftX, ops["kx"], ops["Out"])
    for form in FORMS:
        a8 = _with_form(form, lambda: _c8_inverse(e2, np.asfortranarray(e2[..., 3:]), ops, leftX))
        say(f"  round trip of the beam, mixed {form} per mode (complex64 input): " +
            " ".join(f"{x:.2e}" for x in _per_mode(a8, ref3)))


@pytest.mark.timeout(900)
def test_cfg2_operators(gpu, port, ops):
    """fb_rot (vector -> vector, D+ and D- of modes m+-1) and fb_div (vector
    -> scalar) on FB-space fields of the cfg2 shape."""
    r = np.random.default_rng(2027)
    M = NKO + 1
    v = _field(r, np.exp(-np.arange(NKR) / 64.0), (M, 3))     # a spectrum decaying with the Bessel index
    z3 = np.zeros((NX, NKR, M, 3), complex, order="F")
    z1 = np.zeros((NX, NKR, M), complex, order="F")
    for name, op, out in (("rot", 0, z3), ("div", 2, z1)):
        say(f"case: fb_{name} Nx=8192 Nkr=512 m=0..3")
        ref = getattr(port, "fb_" + name)(out.copy(order="F"), v, ops["Dp"], ops["Dm"], ops["kx"])
        a64 = getattr(gpu, "fb_" + name)(out.copy(order="F"), v, ops["Dp"], ops["Dm"], ops["kx"])
        e64 = _per_mode(a64, ref)
        say("  fp64 per mode: " + " ".join(f"{x:.2e}" for x in e64))
        assert max(e64) < 1e-12
        for form in FORMS:
            a8 = _with_form(form, lambda: _c8_op(op, out.shape, v, ops))
            e8 = _per_mode(a8, ref)
            say(f"  mixed {form} per mode: " + " ".join(f"{x:.2e}" for x in e8))
            assert max(e8) < 1e-5, (form, e8)
