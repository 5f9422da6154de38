"""The StaticKick field update (chimera_main.py:116-128, solvers.py:333-405)
on a resident bunch against the oracle driver: a relativistic bunch in an
analytic undulator whose self-field is rebuilt every step from its
half-step charge density and current, boosted with its mean momentum
(poiss_corr_stat, the static PSATD push maxwell_init_push, field_drift), in a
window that moves with it.

Bars after 1 and 8 steps: FP64 1e-12 and 1e-9 (order of the deposit atomics
and of the mean-momentum sums); mixed precision 1e-4 (float32 fields and
momenta).
"""
import sys

import numpy as np
import pytest

from tests._cases import rel_l2

pytestmark = pytest.mark.gpu

TOL = {"fp64": (1e-12, 1e-9), "mixed": (1e-4, 1e-4)}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _spec(precision):
    from chimera_amd import workloads as W
    spec = W.beam(Nx=128, Nkr=32, gamma=100.0, sig_x=2.0, sig_r=1.0, emittance=0.5, n_target=2e4, K=2.0,
                  lambda_u=150.0, dens=1e-3, steps=8)
    spec["solver"]["Features"] = ("StaticKick",)
    return W.with_precision(spec, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_static_kick_bunch(precision):
    import chimera_amd as G
    from chimera_amd import workloads as W
    from oracle import driver as O
    say(f"case: StaticKick bunch 128x32, m=0,1, {precision}, 8 steps vs oracle")
    spec = _spec(precision)
    g, c = W.build_beam(G, spec, 9), W.build_beam(O, _spec("fp64"), 9)
    assert any("StaticKick" in b for b in g.graph_blockers())
    for i in range(1, 9):
        g.make_step(i)
        c.make_step(i)
        if i not in (1, 8):
            continue
        gs, gp, cs, cp = g.Solvers[0], g.Particles[0], c.Solvers[0], c.Particles[0]
        tol = TOL[precision][0 if i == 1 else 1]
        assert gp.np == cp.Data["coords"].shape[1]
        assert float(gs.Args["leftX"]) == float(cs.Args["leftX"])
        err = {"EG_fb": rel_l2(gs.Data["EG_fb"].numpy().astype(np.complex128), cs.Data["EG_fb"]),
               "EB": rel_l2(gs.Data["EB"].numpy().astype(np.complex128), cs.Data["EB"]),
               "momenta": rel_l2(gp.host("momenta").astype(np.float64), cp.Data["momenta"]),
               "coords": rel_l2(gp.host("coords"), cp.Data["coords"])}
        say(f"  step {i}: " + " ".join(f"{k}={v:.2e}" for k, v in err.items()))
        assert np.abs(cs.Data["EG_fb"]).max() > 0
        for k, v in err.items():
            assert v <= tol, (i, k, v)


def test_momentum_mean():
    """chimera_momentum_mean(_mixed) against the NumPy quotient of the
    reference's expression, on the particles of the bunch; NaN for none."""
    import chimera_amd as G
    from chimera_amd import workloads as W
    say("case: momentum_mean on the bunch particles, fp64 and mixed, and np = 0")
    for precision, tol in (("fp64", 1e-14), ("mixed", 1e-6)):
        sp = W.build_beam(G, _spec(precision), 9).Particles[0]
        u, w = sp.host("momenta").astype(np.float64), sp.host("weights").astype(np.float64)
        ref = (u[0] * w).sum() / w.sum()
        got = sp.momentum_mean()
        assert abs(got - ref) <= tol * abs(ref), (precision, got, ref)
        n = sp.np
        sp.np = 0
        assert np.isnan(sp.momentum_mean())
        sp.np = n
