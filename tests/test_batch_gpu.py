"""Independent simulations on library contexts, recorded and replayed steps,
and batches (chimera_amd.batch).

* A bunch in an undulator with space charge and a window that moves every
  step, stepped through recorded graphs, against the CPU oracle driver
  (FP64: fields and moments <= 1e-9 after 12 steps).
* Recorded replays against eager steps of the same run: the same kernels on
  the same inputs, so the results differ only by the order of the atomic adds
  that the deposition falls back to for tiles it cannot keep in a slot.  Bars:
  1e-12 in FP64; 1e-4 in mixed precision, where that order changes sums by
  FP32 round-off (6e-8) and ten steps of a bunch with space charge amplify it
  about a hundredfold (two eager runs differ as much: printed beside).  The
  device frame follows the host's window exactly.
* A simulation alone and inside a batch of other simulations: each result is
  the same within the same bars.
* Error paths of the context and capture entry points.
"""
import ctypes
import sys

import numpy as np
import pytest

from tests._cases import rel_l2

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "mixed": 1e-4}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_beam(**kw):
    from chimera_amd import workloads as W
    args = dict(Nx=128, Nkr=32, gamma=100.0, sig_x=2.0, sig_r=1.0, emittance=0.5, n_target=2e4, K=2.0,
                lambda_u=150.0, dens=1e-3, steps=12)
    args.update(kw)
    return W.beam(**args)


def snap(run):
    s, e = run.Solvers[0], run.Particles[0]
    return {"EB": s.Data["EB"].numpy().astype(np.complex128), "EG_fb": s.Data["EG_fb"].numpy().astype(np.complex128),
            "momenta": e.host("momenta").astype(np.float64), "coords": e.host("coords"),
            "weights": e.host("weights").astype(np.float64), "np": e.np, "leftX": float(s.Args["leftX"])}


def moments(u, w):
    W = w.sum()
    mean = (u * w).sum(1) / W
    rms = np.sqrt(((u - mean[:, None]) ** 2 * w).sum(1) / W)
    return W, mean, rms


def moment_err(a, b):
    Wa, ma, ra = moments(a["momenta"], a["weights"])
    Wb, mb, rb = moments(b["momenta"], b["weights"])
    return max(abs(Wa - Wb) / abs(Wb), float(np.max(np.abs(ma - mb) / rb)), float(np.max(np.abs(ra - rb) / rb)))


def test_beam_recorded_steps_against_oracle():
    import chimera_amd as G
    from chimera_amd import workloads as W
    from chimera_amd.batch import Simulation
    from oracle import driver as O
    spec = small_beam()
    say("case: beam 128x32, m=0,1, undulator, staged window; recorded steps vs oracle, 12 steps")
    sim = Simulation(lambda: W.build_beam(G, spec, 7), steps=12)
    assert sim.graphs_on, sim.run.graph_blockers()
    sim.advance(12)
    g = snap(sim.result())
    assert set(sim.graphs) == {0, 1}
    c = W.build_beam(O, spec, 7)
    for i in range(1, 13):
        c.make_step(i)
    cs, ce = c.Solvers[0], c.Particles[0]
    ref = {"EB": cs.Data["EB"], "EG_fb": cs.Data["EG_fb"], "momenta": ce.Data["momenta"],
           "weights": ce.Data["weights"], "coords": ce.Data["coords"]}
    out = {k: rel_l2(g[k], ref[k]) for k in ("EB", "EG_fb", "momenta", "coords")}
    out["moments"] = moment_err(g, ref)
    say("  " + " ".join(f"{k}={v:.2e}" for k, v in out.items()))
    assert g["np"] == ce.Data["coords"].shape[1]
    assert g["leftX"] == float(cs.Args["leftX"])
    for k in ("EB", "EG_fb", "moments"):
        assert out[k] <= 1e-9, (k, out[k])
    assert out["momenta"] <= 1e-7 and out["coords"] <= 1e-7
    sim.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_replay_matches_eager(precision):
    import chimera_amd as G
    from chimera_amd import workloads as W
    from chimera_amd.batch import Simulation
    spec = W.with_precision(small_beam(Nx=256, Nkr=48, n_target=5e4), precision)
    say(f"case: beam 256x48 {precision}, recorded vs eager, 10 steps")
    a = Simulation(lambda: W.build_beam(G, spec, 3), steps=10, graphs=True)
    b = Simulation(lambda: W.build_beam(G, spec, 3), steps=10, graphs=False)
    c = Simulation(lambda: W.build_beam(G, spec, 3), steps=10, graphs=False)
    for s in (a, b, c):
        s.advance(10)
    sa, sb, sc = snap(a.result()), snap(b.result()), snap(c.result())
    frame = a.run.Solvers[0].frame.numpy()
    assert frame[1] == sa["leftX"] and sa["leftX"] == sb["leftX"], (frame, sa["leftX"], sb["leftX"])
    tol = TOL[precision]
    for k in ("EB", "EG_fb", "momenta", "coords"):
        e = rel_l2(sa[k], sb[k])
        say(f"  {k}: recorded vs eager {e:.2e}, eager vs eager {rel_l2(sc[k], sb[k]):.2e}")
        assert e <= tol, (k, e)
    for s in (a, b, c):
        s.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_alone_and_in_a_batch(precision):
    import chimera_amd as G
    from chimera_amd import workloads as W
    from chimera_amd.batch import Simulation, run_batch
    specs = [W.with_precision(small_beam(Nx=nx, Nkr=nkr, n_target=n, gamma=gm, steps=8), precision)
             for nx, nkr, n, gm in ((256, 40, 3e4, 80.0), (128, 64, 8e4, 300.0), (512, 32, 2e4, 60.0),
                                    (256, 96, 1e5, 450.0))]
    say(f"case: 4 beams {precision}, alone vs in a batch, 8 steps")
    alone = Simulation(lambda: W.build_beam(G, specs[0], 11), steps=8)
    alone.advance(8)
    s_alone = snap(alone.result())
    alone.close()
    sims = [Simulation(lambda sp=sp, k=k: W.build_beam(G, sp, 11 + k), steps=8) for k, sp in enumerate(specs)]
    run_batch(sims)
    s_mix = snap(sims[0].run)
    tol = TOL[precision]
    for k in ("EB", "EG_fb", "momenta", "coords"):
        e = rel_l2(s_mix[k], s_alone[k])
        say(f"  {k}: {e:.2e}")
        assert e <= tol, (k, e)
    for s in sims:
        assert s.istep == 8 and s.done()
        s.close()


def test_context_and_capture_errors():
    from chimera_amd._lib import lib
    from chimera_amd.batch import Context
    from chimera_amd.device import DeviceArray
    L = lib()
    h = ctypes.c_void_p(0)
    assert L.chimera_capture_end(ctypes.byref(h)) != 0            # not capturing
    assert L.chimera_graph_launch(None) != 0
    ctx = Context()
    with ctx:
        buf = DeviceArray((256,), np.float64)
        assert L.chimera_capture_begin() == 0
        assert L.chimera_capture_begin() != 0                      # already capturing
        assert L.chimera_context_bind(None) != 0                   # cannot leave a capturing context
        assert L.chimera_memset(buf.ptr, 0x3F, buf.nbytes) == 0   # recorded, not run
        assert L.chimera_capture_end(ctypes.byref(h)) == 0
        assert not np.any(buf.numpy())
        assert L.chimera_graph_launch(h.value) == 0
        assert np.all(buf.numpy().view(np.uint8) == 0x3F)
    assert L.chimera_context_destroy(ctx.handle) != 0              # a graph is still alive
    assert L.chimera_graph_destroy(h.value) == 0
    assert ctx.done()
    with ctx:
        del buf
    ctx.close()
    f, t = ctypes.c_int64(0), ctypes.c_int64(0)
    assert L.chimera_mem_info(ctypes.byref(f), ctypes.byref(t)) == 0
    assert 0 < f.value <= t.value


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_served_results_match_alone(precision):
    """Requests arriving while others run (continuous batching, builder thread,
    admission by free memory) give the results each gives alone."""
    import time
    from chimera_amd import serve as S
    specs = [small_beam(Nx=nx, Nkr=nkr, n_target=n, gamma=gm, steps=st)
             for nx, nkr, n, gm, st in ((256, 40, 3e4, 80.0, 10), (128, 64, 8e4, 300.0, 14), (512, 32, 2e4, 60.0, 8),
                                        (256, 96, 1e5, 450.0, 12), (128, 32, 1e4, 120.0, 16))]
    say(f"case: 5 beams {precision} served with staggered arrivals vs alone")
    reqs = [S.Request(f"r{k}", sp, 50 + k, sp["steps"], precision, t_arrival=0.05 * k) for k, sp in enumerate(specs)]
    alone = {}
    for r in reqs:
        sim = r.build()
        sim.advance(r.steps)
        alone[r.name] = snap(sim.result())
        sim.close()
    served = {}
    server = S.Server(slots=3, reserve=0, chunk=2, on_finish=lambda r: served.__setitem__(r.name, snap(r.sim.run)))
    try:
        S.serve(reqs, server)
    finally:
        server.close()
    assert sorted(served) == sorted(alone)
    assert all(r.t_done >= r.t_arrival_host for r in server.finished)
    assert all(r.result is not None and len(r.result) == 1 and r.result[0].shape == (3, 3) for r in server.finished)
    for name in alone:
        for k in ("EB", "EG_fb", "momenta", "coords"):
            e = rel_l2(served[name][k], alone[name][k])
            assert e <= TOL[precision], (name, k, e)
    say("  all requests match their runs alone")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_evicted_requests_resume(precision):
    """With device memory for one request at a time (an emulated budget) and
    eviction on, later arrivals evict the running request, whose state goes to
    host memory and resumes from there; every request still gives the result
    it gives alone."""
    from chimera_amd import serve as S
    # one grid and particle count (the same memory estimate), three bunches
    specs = [small_beam(Nx=256, Nkr=40, n_target=3e4, gamma=gm, steps=st) for gm, st in ((80.0, 90), (300.0, 80), (60.0, 70))]
    say(f"case: 3 beams {precision}, memory for one at a time, eviction on, vs alone")
    # all arrive at once: the first runs, the others wait for memory and evict it
    reqs = [S.Request(f"e{k}", sp, 70 + k, sp["steps"], precision, t_arrival=0.0) for k, sp in enumerate(specs)]
    alone = {}
    for r in reqs:
        sim = r.build()
        sim.advance(r.steps)
        alone[r.name] = snap(sim.result())
        sim.close()
    served = {}
    assert len({r.bytes for r in reqs}) == 1
    budget = int(1.5 * reqs[0].bytes)             # one request at a time
    server = S.Server(slots=3, reserve=0, chunk=2, evict=True, hbm=budget, log=say,
                      on_finish=lambda r: served.__setitem__(r.name, snap(r.sim.run)))
    try:
        S.serve(reqs, server)
    finally:
        server.close()
    assert sorted(served) == sorted(alone)
    n_evict = sum(r.evictions for r in server.finished)
    say(f"  {n_evict} evictions")
    assert n_evict >= 1
    for name in alone:
        for k in ("EB", "EG_fb", "momenta", "coords"):
            e = rel_l2(served[name][k], alone[name][k])
            assert e <= TOL[precision], (name, k, e)
    say("  every request matches its run alone")
