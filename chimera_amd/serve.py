"""Scans and a request loop of independent PIC simulations on one GPU.

The reference runs one simulation per process; this module keeps many
resident at once (chimera_amd.batch: one library context and stream each,
steps recorded into graphs and replayed) and schedules them:

* run_scan: a fixed set of simulations (a parameter scan) to completion,
  either one after another with eager steps (the plain path) or all at once
  with their steps interleaved (the batched path); reports particle-steps per
  second and the time-to-result of every simulation.
* Server: continuous batching.  Requests arrive over time; a request is
  admitted when a slot is free and its estimated device memory fits in the
  free HBM (less a reserve); it is built (plasma generation, tables, initial
  half step) on one of a few builder threads, then advanced a few steps at a time in turn
  with every other active simulation, so that the work of requests that
  arrived at different times overlaps on the device.  A finished
  simulation's result, the beam envelopes of its pushed species, is reduced
  on the device into Request.result; then its arrays and context are freed.
  With evict=True a request that does not fit may evict the active request
  with the most steps left: its state goes to host memory (the arrays of
  ChimeraRun.snapshot, with the request's own random stream), its device
  memory is freed, and it queues again to resume from there.

A request ends in one of three states (Request.status): "done" with its
result; "failed" (its build raised, e.g. an allocation failure, or its result
holds non-finite values, e.g. particles that went NaN), with the reason in
Request.error; "cancelled" (Server.cancel).  A failed or cancelled request
frees its slot and memory and leaves the others running.
"""
from __future__ import annotations

import queue
import threading
import time

import numpy as np

from .batch import Simulation, free_device_bytes
from .diagnostics import Diagnostics

GIB = 1 << 30


def estimate_bytes(spec, precision="fp64"):
    """Device memory a workload needs, from its grid and particle counts, with a
    1.5x margin for capacity growth and scratch memory."""
    s = spec["solver"]
    xmin, xmax, rmax, dx, dr = s["Grid"]
    Nx = int(round((xmax - xmin) / dx))
    Nkr = int(round(rmax / dr))
    M = int(s["MaxAzimuthMode"]) + 1
    cf = 16 if precision == "fp64" else 8
    su = 8 if precision == "fp64" else 4
    grid = Nx * (Nkr + 1) * M * cf * (6 + 3 + 2)              # EB, J, rho (+ background)
    fb = Nx * Nkr * M * cf * (6 + 3 + 3 + 1 + 6 + 3 + 6)       # FB arrays + transform scratch
    tables = Nx * Nkr * M * 10 * su + 4 * Nkr * Nkr * (M + 1) * su
    if spec.get("kind") == "beam":
        n = spec["params"]["n_target"] * 1.3
        parts = n * (48 + 4 * su)
    else:
        fc = spec["electrons"]["FixedCell"]
        n = Nx * Nkr * int(np.prod(fc))
        parts = 2 * n * (48 + 4 * su)
    return int(1.5 * (grid + fb + tables + parts))


class Request:
    """One simulation to run: a workload spec, its seed, precision and step count."""

    def __init__(self, name, spec, seed, steps, precision="mixed", t_arrival=0.0):
        from . import workloads as W
        self.name, self.seed, self.steps, self.precision = name, int(seed), int(steps), precision
        self.spec = W.with_precision(spec, precision)
        self.t_arrival = float(t_arrival)
        self.bytes = estimate_bytes(self.spec, precision)
        self.sim = None
        self.t_admit = self.t_built = self.t_done = None
        self.particle_steps = 0
        self.result = None             # beam envelopes of the finished run (diagnostics.get_beam_envelops)
        self.snap, self.istep0, self.evictions = None, 0, 0     # state of an evicted run
        self.status, self.error = "pending", None
        self.cancel_requested = False

    def build(self, graphs=True):
        import chimera_amd as G
        from . import workloads as W
        if self.spec.get("kind") == "beam":
            fn = lambda: W.build_beam(G, self.spec, self.seed, own_rng=True)
        else:
            fn = lambda: W.build(G, self.spec, self.seed, own_rng=True)
        if self.snap is not None:
            def resume(make=fn, snap=self.snap):
                run = make()
                run.read_snap(snap)
                return run
            sim = Simulation(resume, self.steps, graphs=graphs, name=self.name, start=self.istep0)
            self.snap = None
            return sim
        return Simulation(fn, self.steps, graphs=graphs, name=self.name)


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=float), q)) if len(xs) else None


def run_scan(requests, batched=True, graphs=True, chunk=4, gate=True, log=None):
    """Run every request to completion and return (particle-steps/s over the
    stepping, per-request time-to-result in s measured from the start of the
    stepping, stepping seconds).  Building is not timed.  batched=False runs
    the simulations one after another (each built, stepped and freed before
    the next); batched=True builds them all, then interleaves their steps,
    `chunk` steps of each in turn (with gate, a simulation's next chunk waits
    until its previous one has finished on the device)."""
    if not batched:
        psteps, wall, ttr = 0, 0.0, []
        for r in requests:
            sim = r.build(graphs=graphs)
            t0 = time.perf_counter()
            sim.advance(r.steps)
            sim.result()
            dt = time.perf_counter() - t0
            wall += dt
            ttr.append(wall)
            psteps += sim.particles * r.steps
            r.particle_steps = sim.particles * r.steps
            sim.close()
            if log:
                log(f"  {r.name}: {sim.particles} particles, {r.steps} steps in {dt * 1e3:.1f} ms")
        return psteps / wall, ttr, wall
    sims = [r.build(graphs=graphs) for r in requests]
    for r, s in zip(requests, sims):
        r.particle_steps = s.particles * r.steps
    # shortest first: small simulations reach the device first and finish early
    order = sorted(range(len(sims)), key=lambda k: requests[k].particle_steps)
    t0 = time.perf_counter()
    done = {}
    while len(done) < len(sims):
        for k in order:
            if k in done:
                continue
            s = sims[k]
            if s.finished_issuing:
                if s.done():
                    done[k] = s.t_done - t0
            elif not gate or s.istep == 0 or s.ctx.done():
                s.advance(chunk)
    wall = time.perf_counter() - t0
    psteps = sum(r.particle_steps for r in requests)
    for s in sims:
        s.close()
    return psteps / wall, [done[k] for k in range(len(sims))], wall


class Server:
    """Continuous batching of requests on one GPU (see the module docstring).

    slots: most simulations resident at once; reserve: device bytes kept free;
    chunk: steps issued per simulation per turn (a simulation gets its next
    chunk once the previous one has finished)."""

    def __init__(self, slots=64, reserve=16 * GIB, chunk=4, graphs=True, builder_thread=True, log=None,
                 on_finish=None, builders=4, evict=False, hbm=None, max_evictions=2, evict_after=16,
                 free_bytes=None):
        self.slots, self.reserve, self.chunk, self.graphs = slots, reserve, chunk, graphs
        # free_bytes: () -> free device bytes (default: the device's, chimera_mem_info)
        self.free_bytes = free_bytes if free_bytes is not None else (lambda: free_device_bytes()[0])
        # evict: a request that does not fit may evict the active request with the most
        # steps left once it has run evict_after steps since its admission (its state to
        # host memory, resumed later from there; at most max_evictions times); hbm: device
        # bytes the server may use (default: whatever the device has free)
        self.evict, self.hbm, self.max_evictions, self.evict_after = evict, hbm, max_evictions, evict_after
        self.on_finish = on_finish     # called with each finished request before its simulation is freed
        self.log = log
        self.pending = []          # arrived, not admitted (FIFO)
        self.building = 0
        self.reserved = 0          # estimated bytes of admitted requests not yet allocated
        self.active = []
        self.finished = []         # done, in order of completion
        self.failed = []           # failed or cancelled
        self._in_build = set()     # admitted, being built
        self.built = queue.Queue()
        self.errors = []           # (name, reason) of every failed request
        self.builder_thread = builder_thread
        if builder_thread:
            # several builders: plasma generation and table construction are
            # NumPy/SciPy work that runs outside the interpreter lock
            self.jobs = queue.Queue()
            self.workers = [threading.Thread(target=self._builder, daemon=True) for _ in range(max(1, builders))]
            for w in self.workers:
                w.start()

    def _builder(self):
        while True:
            r = self.jobs.get()
            if r is None:
                return
            self._build(r)
            self.built.put(r)

    def _build(self, r):
        try:
            r.sim = r.build(graphs=self.graphs)
            r.t_built = time.perf_counter()
        except Exception as e:          # the request fails; the loop goes on
            r.sim = None
            r.error = f"build failed: {e!r}"

    def submit(self, r, now):
        r.t_arrival_host = now
        self.pending.append(r)

    def _fail(self, r, reason, status="failed"):
        r.status, r.error = status, reason
        r.t_done = time.perf_counter()
        self.failed.append(r)
        self.errors.append((r.name, reason))
        if self.log:
            self.log(f"  {r.name} {status}: {reason}")

    def cancel(self, name):
        """Cancel a request by name: a pending one leaves the queue, an active
        one is stopped and freed (its work in flight finishes first), one being
        built is dropped once built.  Returns whether a request was found."""
        for r in list(self.pending):
            if r.name == name:
                self.pending.remove(r)
                self._fail(r, "cancelled", "cancelled")
                return True
        for r in list(self.active):
            if r.name == name:
                self.active.remove(r)
                if r.sim is not None:
                    r.sim.close()
                    r.sim = None
                self._fail(r, "cancelled", "cancelled")
                return True
        r = next((r for r in self._in_build if r.name == name), None)
        if r is not None:
            r.cancel_requested = True
            return True
        return False

    def result_of(self, sim):
        """The result of a finished simulation: the beam envelopes of its pushed
        species, reduced on the device."""
        with sim.ctx:
            return Diagnostics(sim.run).get_beam_envelops()

    def _free(self):
        free = self.free_bytes()
        if self.hbm is not None:
            free = min(free, self.hbm - sum(a.bytes for a in self.active))
        return free

    def _evict_for(self, r, free):
        """Evict the active request with the most steps left whose memory makes
        room for r, if it has run evict_after steps since it was admitted and
        has not been evicted too often; returns whether one was evicted."""
        need = r.bytes + self.reserve + self.reserved - free
        cands = [a for a in self.active if a.sim is not None and not a.sim.finished_issuing
                 and a.sim.istep >= a.istep0 + max(1, self.evict_after) and a.evictions < self.max_evictions
                 and a.bytes >= need]
        if not cands:
            return False
        v = max(cands, key=lambda a: a.sim.steps - a.sim.istep)
        s = v.sim
        s.ctx.synchronize()
        with s.ctx:
            v.snap = s.run.snapshot()
        v.istep0 = s.istep
        v.evictions += 1
        s.close()
        v.sim = None
        self.active.remove(v)
        self.pending.append(v)
        if self.log:
            self.log(f"  evicted {v.name} after step {v.istep0} for {r.name}")
        return True

    def _admit(self, now):
        free = self._free()
        while self.pending and len(self.active) + self.building < self.slots:
            r = self.pending[0]
            if r.bytes + self.reserve + self.reserved > free:
                if self.evict and self._evict_for(r, free):
                    free = self._free()
                    continue
                break
            self.pending.pop(0)
            r.t_admit = now
            r.status = "building"
            self.reserved += r.bytes
            self.building += 1
            self._in_build.add(r)
            if self.builder_thread:
                self.jobs.put(r)
            else:
                self._build(r)
                self.built.put(r)

    def poll(self, now):
        """One turn: take in built simulations, issue work, retire finished ones, admit."""
        while not self.built.empty():
            r = self.built.get()
            self.building -= 1
            self._in_build.discard(r)
            self.reserved -= r.bytes        # now allocated: free_device_bytes() accounts for it
            if r.sim is None:
                self._fail(r, r.error or "build failed")
                continue
            if r.cancel_requested:
                r.sim.close()
                r.sim = None
                self._fail(r, "cancelled", "cancelled")
                continue
            r.particle_steps = r.sim.particles * r.steps
            r.status = "running"
            self.active.append(r)
        still = []
        for r in self.active:
            s = r.sim
            if s.finished_issuing:
                if s.done():
                    r.t_done = s.t_done
                    r.result = self.result_of(s)
                    if self.on_finish is not None:
                        self.on_finish(r)
                    s.close()
                    r.sim = None
                    bad = [k for k, v in enumerate(r.result or ()) if not np.all(np.isfinite(np.asarray(v)))]
                    if bad:
                        self._fail(r, f"non-finite result (species {bad})")
                    else:
                        r.status = "done"
                        self.finished.append(r)
                    continue
            elif s.istep == 0 or s.ctx.done():
                s.advance(self.chunk)
            still.append(r)
        self.active = still
        self._admit(now)

    @property
    def idle(self):
        return not (self.pending or self.active or self.building)

    def close(self):
        if self.builder_thread:
            for _ in self.workers:
                self.jobs.put(None)
            for w in self.workers:
                w.join()


MIX = {"lpa_cfg0": 0.45, "beam": 0.5, "lpa_cfg1": 0.05}


def poisson_requests(rate, seconds, seed, precision="mixed", mix=None):
    """Requests arriving as a Poisson stream of `rate` per second for `seconds`,
    each drawn from the mix of workloads (weights per kind) with scan
    parameters from one seed: plasma density 1e-4..1e-2 n_c (log-uniform) and
    a0 0.5..3 for the laser-plasma kinds; beam charge (density 1e-4..1e-2,
    log-uniform), emittance and the other bunch parameters of
    workloads.beam_scan for the bunches."""
    from . import workloads as W
    mix = MIX if mix is None else mix
    kinds = sorted(mix)
    p = np.array([mix[k] for k in kinds], dtype=float)
    p /= p.sum()
    rng = np.random.RandomState(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > seconds:
            return out
        kind = kinds[rng.choice(len(kinds), p=p)]
        if kind == "beam":
            spec = W.beam_scan(1, seed=int(rng.randint(1 << 30)), steps=300)[0]
            spec["beam"]["Density"] = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2))))
        else:
            dens = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2))))
            a0 = float(rng.uniform(0.5, 3.0))
            spec = W.WORKLOADS[kind](dens=dens, a0=a0)
        out.append(Request(f"{kind}#{len(out)}", spec, int(rng.randint(1 << 30)), spec["steps"], precision, t))


def serve(requests, server, log=None, t0=None):
    """Feed requests (sorted by t_arrival, seconds after start) to the server in
    real time and run until all have finished; returns the wall time."""
    t0 = time.perf_counter() if t0 is None else t0
    i = 0
    last = t0
    while i < len(requests) or not server.idle:
        now = time.perf_counter()
        while i < len(requests) and requests[i].t_arrival <= now - t0:
            server.submit(requests[i], now)
            i += 1
        server.poll(now)
        if log and now - last > 5.0:
            last = now
            log(f"  t={now - t0:6.1f}s arrived {i} pending {len(server.pending)} building {server.building} "
                f"active {len(server.active)} finished {len(server.finished)}")
        if not server.active and not server.built.qsize() and i < len(requests) and not server.pending:
            time.sleep(min(0.002, max(0.0, requests[i].t_arrival - (time.perf_counter() - t0))))
    return time.perf_counter() - t0
