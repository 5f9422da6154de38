"""The benchmark configurations as whole runs on the GPU against the FP64
oracle driver, from identical seeded inputs, with the oracle's own noise floor
(the same oracle run with its particles shuffled) printed beside each
comparison.

* lpa_cfg1 (4096 x 256, modes 0..2, 16.7 M electrons, moving window with
  absorption and injection every 10 steps): half step and step 1 at relative
  L2 <= 1e-12 for EG_fb and EB and particle by particle; then to step 11,
  across one window move, at <= 1e-9 for the fields and the particle moments.
  J_fb after step 1 is held to max(1e-12, 4 x the oracle's rounding floor):
  the oracle compiled with FMA contraction (the same arithmetic, rounded once
  per multiply-add) differs from itself there by about 7e-12 at this size,
  because the Poisson correction's (grad rho_n+1 - grad rho_n)/dt cancels.
* lpa_cfg2 (8192 x 512, modes 0..3, 67 M electrons): half step and step 1 at
  <= 1e-12, J_fb after step 1 again against the oracle's rounding floor.
* The bunch of the batch workloads (space charge, analytic undulator, window
  following the bunch) in mixed precision against the FP64 oracle: real-space
  fields and particle moments <= 1e-5 after 12 steps.
* Served requests that include lpa_cfg0-shaped runs and a run whose moving
  window injects plasma, with eviction: every result equals its run alone.
"""
import sys

import numpy as np
import pytest

from tests._cases import rel_l2
from tests.test_step_gpu import (compare_snaps, moments, moments_err, noise_floor, rounding_floor, snapshot_cpu,
                                 snapshot_gpu)

pytestmark = pytest.mark.gpu


_CAPTURE = None


@pytest.fixture(autouse=True)
def _progress_uncaptured(capsys):
    """These runs take minutes: their progress lines go out as they happen,
    past pytest's output capture, so that a quiet stretch is never mistaken
    for a hang."""
    global _CAPTURE
    _CAPTURE = capsys
    yield
    _CAPTURE = None


def say(msg):
    if _CAPTURE is not None:
        with _CAPTURE.disabled():
            print(msg, flush=True)
    else:
        print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.timeout(1200)
def test_lpa_cfg1_against_oracle():
    import chimera_amd as G
    from chimera_amd import workloads as W
    from oracle import driver as O
    say("case: lpa_cfg1 half step, step 1, step 11 (one window move)")
    spec = W.lpa_cfg1()
    g = W.build(G, spec, 41)
    assert (g.Solvers[0].Args["Nx"], g.Solvers[0].Args["Nkr"]) == (4096, 256)
    snaps = {0: snapshot_gpu(g)}
    for i in range(1, 12):
        g.make_step(i)
        if i in (1, 10, 11):
            snaps[i] = snapshot_gpu(g)
    say(f"  GPU run done ({g.Particles[0].np} electrons at step 11); oracle runs")
    # the shuffled run first: the injected plasma is drawn from NumPy's global
    # stream, so each oracle run goes to the end on its own after its seeding
    cs = W.build(O, spec, 41)
    cs.shuffle_particles(5)
    cs.make_step(1)
    floor1 = snapshot_cpu(cs)
    del cs
    say("  shuffled oracle run: step 1 done")
    with O.rounding_variant():
        cf = W.build(O, spec, 41)
        cf.make_step(1)
    fma1 = snapshot_cpu(cf)
    del cf
    say("  FMA-contracted oracle run: step 1 done")
    c = W.build(O, spec, 41)
    compare_snaps(snaps[0], snapshot_cpu(c), 1e-12, 1e-12, "  cfg1 half step")
    for i in range(1, 12):
        c.make_step(i)
        say(f"  oracle step {i}")
        if i == 1:
            s1 = snapshot_cpu(c)
            noise_floor(s1, floor1, "  cfg1 step 1")
            rf = rounding_floor(s1, fma1, "  cfg1 step 1")
            assert max(rf["EG_fb"], rf["EB"]) < 1e-13      # the fields themselves are well conditioned
            compare_snaps(snaps[1], s1, 1e-12, 1e-12, "  cfg1 step 1", floor={"EG_fb": 0, "EB": 0, "J_fb": rf["J_fb"]})
        elif i in (10, 11):
            cs_ = snapshot_cpu(c)
            out = {k: rel_l2(snaps[i][k], cs_[k]) for k in ("EG_fb", "J_fb", "EB")}
            out["moments"] = moments_err(moments(snaps[i]["momenta"], snaps[i]["weights"]),
                                         moments(cs_["momenta"], cs_["weights"]))
            out["momenta"] = rel_l2(snaps[i]["momenta"], cs_["momenta"])
            say(f"  cfg1 step {i}: " + " ".join(f"{k}={v:.2e}" for k, v in out.items()))
            assert snaps[i]["np"] == cs_["np"], (i, snaps[i]["np"], cs_["np"])
            for k in ("EG_fb", "J_fb", "EB", "moments"):
                assert out[k] <= 1e-9, (i, k, out[k])
            assert out["momenta"] <= 1e-7, (i, out["momenta"])


@pytest.mark.timeout(1200)
def test_lpa_cfg2_against_oracle():
    import chimera_amd as G
    from chimera_amd import workloads as W
    from oracle import driver as O
    say("case: lpa_cfg2 half step and step 1")
    spec = W.lpa_cfg2()
    g = W.build(G, spec, 43)
    assert (g.Solvers[0].Args["Nx"], g.Solvers[0].Args["Nkr"], g.Solvers[0].Args["Mtot"]) == (8192, 512, 4)
    s0 = snapshot_gpu(g)
    g.make_step(1)
    s1 = snapshot_gpu(g)
    np_gpu = g.Particles[0].np
    del g
    say(f"  GPU run done ({np_gpu} electrons); oracle runs")
    with O.rounding_variant():
        cf = W.build(O, spec, 43)
        say("  FMA-contracted oracle run: built, half step done")
        cf.make_step(1)
    fma1 = snapshot_cpu(cf)
    del cf
    say("  FMA-contracted oracle run: step 1 done")
    c = W.build(O, spec, 43)
    say("  oracle run: built, half step done")
    compare_snaps(s0, snapshot_cpu(c), 1e-12, 1e-12, "  cfg2 half step")
    del s0
    c.make_step(1)
    c1 = snapshot_cpu(c)
    rf = rounding_floor(c1, fma1, "  cfg2 step 1")
    assert max(rf["EG_fb"], rf["EB"]) < 1e-13
    compare_snaps(s1, c1, 1e-12, 1e-12, "  cfg2 step 1", floor={"EG_fb": 0, "EB": 0, "J_fb": rf["J_fb"]})


def test_mixed_bunch_against_fp64_oracle():
    """The batch workloads' precision: a bunch with space charge in an
    undulator, mixed precision on the GPU (split-BF16 GEMMs, and the FP32-MFMA
    form beside it), FP64 on the oracle, 12 steps.  Bars: the real-space
    fields EB and the particle moments <= 1e-5.  The FB-space EG_fb is
    printed with the spread between the two mixed forms: with the window
    moving with the bunch the density changes by ~1e-3 of itself per step, so
    the Poisson correction's (grad rho_n+1 - grad rho_n)/dt amplifies the
    complex64 rounding of rho, in every mixed form alike."""
    import chimera_amd as G
    from chimera_amd import workloads as W
    from chimera_amd._lib import lib
    from oracle import driver as O
    say("case: beam 256x40 m=0,1, space charge + undulator, mixed GPU vs FP64 oracle, 12 steps")
    spec = W.beam(Nx=256, Nkr=40, gamma=100.0, sig_x=2.0, sig_r=1.0, emittance=0.5, n_target=3e4, K=2.0,
                  lambda_u=150.0, dens=1e-3, steps=12)
    L = lib()
    old = L.chimera_get_mixed_gemm()

    def gpu_run(form):
        L.chimera_set_mixed_gemm(form)
        try:
            g = W.build_beam(G, W.with_precision(spec, "mixed"), 17)
            for i in range(1, 13):
                g.make_step(i)
        finally:
            L.chimera_set_mixed_gemm(old)
        s, e = g.Solvers[0], g.Particles[0]
        return {"EB": s.Data["EB"].numpy().astype(np.complex128), "EG_fb": s.Data["EG_fb"].numpy().astype(np.complex128),
                "momenta": e.host("momenta").astype(np.float64), "weights": e.host("weights").astype(np.float64)}

    runs = {"bf16x6": gpu_run(0), "fp32": gpu_run(1)}
    c = W.build_beam(O, spec, 17)
    for i in range(1, 13):
        c.make_step(i)
    cs, ce = c.Solvers[0], c.Particles[0]
    mb = moments(ce.Data["momenta"], ce.Data["weights"])
    spread = rel_l2(runs["bf16x6"]["EG_fb"], runs["fp32"]["EG_fb"])
    say(f"  EG_fb spread between the two mixed forms: {spread:.2e}")
    for form, a in runs.items():
        out = {k: rel_l2(a[k], cs.Data[k]) for k in ("EB", "EG_fb")}
        out["moments"] = moments_err(moments(a["momenta"], a["weights"]), mb)
        out["momenta"] = rel_l2(a["momenta"], ce.Data["momenta"])
        say(f"  {form}: " + " ".join(f"{k}={v:.2e}" for k, v in out.items()))
        for k in ("EB", "moments"):
            assert out[k] <= 1e-5, (form, k, out[k])
        assert out["EG_fb"] <= max(1e-5, 4 * spread), (form, out["EG_fb"], spread)


def _snap(run):
    s, e = run.Solvers[0], run.Particles[0]
    return {"EB": s.Data["EB"].numpy(), "EG_fb": s.Data["EG_fb"].numpy(), "momenta": e.host("momenta"),
            "coords": e.host("coords"), "np": e.np}


@pytest.mark.timeout(600)
def test_served_lpa_requests_match_alone():
    """A served mix with lpa_cfg0-shaped requests and a moving-window run that
    injects plasma from its own random stream; device memory for one request
    at a time, so later arrivals evict running ones, whose state (random
    stream included) goes to host memory and resumes from there.  FP64: the
    served results equal the runs alone to 1e-12."""
    from chimera_amd import serve as S
    from chimera_amd import workloads as W
    specs = [("cfg0", W.lpa_cfg0(steps=8), 8), ("inject", W.lpa_small(Nx=256, Nkr=32, nko=1, moving=True, steps=24), 24),
             ("cfg0b", W.lpa_cfg0(steps=6, a0=1.5), 6)]
    say("case: served lpa_cfg0 x2 + injecting moving window, fp64, one at a time with eviction, vs alone")
    reqs = [S.Request(name, sp, 80 + k, st, "fp64", t_arrival=0.0) for k, (name, sp, st) in enumerate(specs)]
    alone = {}
    for r in reqs:
        sim = r.build()
        sim.advance(r.steps)
        alone[r.name] = _snap(sim.result())
        sim.close()
    say("  alone runs done")
    served = {}
    budget = int(1.2 * max(r.bytes for r in reqs))
    server = S.Server(slots=3, reserve=0, chunk=2, evict=True, hbm=budget, log=say,
                      on_finish=lambda r: served.__setitem__(r.name, _snap(r.sim.run)))
    try:
        S.serve(reqs, server)
    finally:
        server.close()
    assert sorted(served) == sorted(alone)
    n_evict = sum(r.evictions for r in server.finished)
    say(f"  {n_evict} evictions")
    assert n_evict >= 1
    for name in alone:
        assert served[name]["np"] == alone[name]["np"], name
        for k in ("EB", "EG_fb", "momenta", "coords"):
            e = rel_l2(served[name][k], alone[name][k])
            assert e <= 1e-12, (name, k, e)
    say("  every request matches its run alone")
