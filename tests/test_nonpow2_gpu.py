"""x transforms of lengths that are not powers of two (Bluestein convolution
in the line kernel), through the C ABI and through the whole resident step,
against the FP64 oracle.

The reference plans an FFTW transform of any length (fb_io.f90:37-38, :51);
its own CI scenario doc/tests/lpa-testrun.py runs at Nx = 528 = 2^4*3*11 and
its LPA demo at Nx = 1272 = 2^3*3*53.  Bars as for the power-of-two cases:
relative L2 <= 1e-12 for one transform, <= 1e-12 after one step and <= 1e-9
after 21 steps (two window moves) for fields and particle moments.
"""
import sys

import numpy as np
import pytest

from tests._cases import C, F, rel_l2, rng

pytestmark = pytest.mark.gpu

TOL_FB = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# 528 and 1272: the reference's scenarios; 24: short lines, many per block;
# 1000: lines of 2048 convolution points; 2500 (odd factor 5^4*4) and 4095
# (odd): 8192 convolution points, one line per workgroup
@pytest.mark.parametrize("nkx,nr,nkr,M", [(528, 16, 16, 2), (1272, 8, 8, 2), (24, 5, 7, 3), (1000, 6, 6, 1),
                                          (2500, 4, 4, 1), (4095, 3, 3, 1)])
def test_fb_transforms_any_length(gpu, port, nkx, nr, nkr, M):
    say(f"case: fb_vec_in/fb_scl_in/fb_eb_out nkx={nkx} nr={nr} nkr={nkr} M={M}")
    r = rng(1700 + nkx)
    vec = C(r, nkx, nr + 1, M, 3)
    In = F(r, nr, nkr, M)
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.04)
    z = np.zeros((nkx, nkr, M, 3), complex, order="F")
    a = gpu.fb_vec_in(z.copy(order="F"), vec, -19.98, kx, In)
    b = port.fb_vec_in(z.copy(order="F"), vec, -19.98, kx, In)
    assert rel_l2(a, b) < TOL_FB
    s = vec[..., 2].copy(order="F")
    a = gpu.fb_scl_in(z[..., 0].copy(order="F"), s, -19.98, kx, In)
    b = port.fb_scl_in(z[..., 0].copy(order="F"), s, -19.98, kx, In)
    assert rel_l2(a, b) < TOL_FB
    e, bb = C(r, nkx, nkr, M, 6), C(r, nkx, nkr, M, 3)
    Out = F(r, nkr, nr, M)
    eb = np.ones((nkx, nr + 1, M, 6), complex, order="F")
    a = gpu.fb_eb_out(eb.copy(order="F"), e, bb, 0.52, kx, Out)
    b = port.fb_eb_out(eb.copy(order="F"), e, bb, 0.52, kx, Out)
    assert rel_l2(a, b) < TOL_FB
    assert not np.any(a[:, 0])


@pytest.mark.parametrize("nkx", [528, 1272])
def test_fb_filtr_any_length(gpu, port, nkx):
    say(f"case: fb_filtr nkx={nkx}")
    r = rng(1800 + nkx)
    v = C(r, nkx, 6, 2, 3)
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.04)
    N = 175
    i = np.arange(N)
    filt = (i >= 0.75 * N) * (0.5 - 0.5 * np.cos(np.pi * (i - 0.75 * N) / (0.25 * N))) ** 2
    a = gpu.fb_filtr(v.copy(order="F"), -19.98, kx, filt, 0)
    b = port.fb_filtr(v.copy(order="F"), -19.98, kx, filt, 0)
    assert rel_l2(a, b) < TOL_FB


def test_fb_length_beyond_the_line_kernel(gpu):
    say("case: fb_vec_in nkx=4097 (refused)")
    nkx, nr, nkr, M = 4097, 2, 2, 1
    vec = np.zeros((nkx, nr + 1, M, 3), complex, order="F")
    In = np.zeros((nr, nkr, M), order="F")
    kx = 2 * np.pi * np.fft.fftfreq(nkx, 0.04)
    with pytest.raises(RuntimeError):
        gpu.fb_vec_in(np.zeros((nkx, nkr, M, 3), complex, order="F"), vec, 0.0, kx, In)


@pytest.mark.timeout(300)
def test_reference_lpa_testrun_21_steps():
    """doc/tests/lpa-testrun.py's dictionaries, unchanged (Nx = 528 on 4
    x-chunks, NoSorting electrons, density ramp injected, 175-point absorbing
    layer): 21 steps, i.e. two window moves, on the GPU and on the oracle."""
    import chimera_amd as G
    from chimera_amd import workloads as W
    from oracle import driver as O
    from tests.test_step_gpu import compare_snaps, noise_floor, snapshot_cpu, snapshot_gpu
    say("case: lpa_testrun (Nx=528) 21 steps")
    spec = W.lpa_testrun()
    checks = (1, 10, 11, 21)
    snaps = {}
    g = W.build(G, spec, 31)
    assert g.Solvers[0].Args["Nx"] == 528
    for i in range(1, 22):
        g.make_step(i)
        if i in checks:
            snaps[i] = snapshot_gpu(g)
    say("  GPU run done; oracle runs (about 0.3 s per step each)")
    # the shuffled run first: the injected plasma is drawn from NumPy's global
    # stream, so each oracle run goes to the end on its own after its seeding
    floor = {}
    c2 = W.build(O, spec, 31)
    c2.shuffle_particles(3)
    for i in range(1, 22):
        c2.make_step(i)
        if i in checks:
            floor[i] = snapshot_cpu(c2)
    c = W.build(O, spec, 31)
    for i in range(1, 22):
        c.make_step(i)
        if i in checks:
            cs = snapshot_cpu(c)
            noise_floor(cs, floor[i], f"step {i}")
            tol = 1e-12 if i == 1 else 1e-9
            compare_snaps(snaps[i], cs, tol, tol, f"step {i}")
