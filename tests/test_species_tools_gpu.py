"""Set-up tools of the species on resident particles: beam_focus
(species.py:473-486) against the reference's formula applied to host copies
of the same particles, in FP64 and mixed precision."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_beam_focus(precision):
    import chimera_amd as G
    from chimera_amd import workloads as W
    spec = W.with_precision(W.beam(Nx=256, Nkr=48, gamma=150.0, sig_x=2.0, sig_r=1.0, emittance=0.5, n_target=5e4,
                                   K=2.0, lambda_u=150.0, steps=4), precision)
    run = W.build_beam(G, spec, 3)
    sp = run.Particles[0]
    x, u, w = sp.host("coords"), sp.host("momenta").astype(np.float64), sp.host("weights").astype(np.float64)
    gg = (1 + u ** 2).sum(0) ** 0.5
    pz = (u[0] * w).sum() / w.sum()
    ref = x.copy()
    ref[0] = ref[0] - 25.0 * (u[0] - pz) / pz / gg ** 2
    ref[1:3] = ref[1:3] - u[1:3] / pz * 25.0
    sp.beam_focus(25.0)
    assert np.array_equal(sp.host("coords"), ref)
    assert np.array_equal(sp.host("coords_halfstep"), ref)
    print(f"case: beam_focus {precision}, {sp.np} particles, max shift {np.abs(ref - x).max():.3e}", flush=True)
    run.make_step(1)        # the run steps on from the focused positions
