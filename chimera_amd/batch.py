"""Many independent simulations on one GPU: each resident in device memory on
its own library context (stream, scratch memory, tables), each step recorded
once into a HIP graph and replayed, and the simulations' steps interleaved so
that their work overlaps on the device.

The reference runs one simulation per process and has no counterpart; the
simulations themselves are ChimeraRun instances (chimera_main.py:23-318)
stepped exactly as make_step does.

    sim = Simulation(lambda: workloads.build_beam(chimera_amd, spec), steps=300)
    run_batch([sim, ...])            # all to completion, round-robin

Recording: steps 1 and 2 run eagerly (they size the scratch memory and upload
the job tables of both arrangements of the swapped density-gradient buffers);
steps 3 and 4 are recorded while they are issued and then launched; from step
5 on, step i replays the graph of its parity after the host-side bookkeeping
of make_step (window moves, buffer swap).  A run resumed from a snapshot after
step s does the same from step s + 1.  A run whose steps differ from one
another (plasma injection, absorption, re-sorting, a window that does not move
every step; ChimeraRun.graph_blockers) is stepped eagerly on its context.
"""
from __future__ import annotations

import ctypes
import time

from ._lib import DEVICE_LOCK, check, lib


class Context:
    """A libchimera_hip context; `with ctx:` makes it current for this thread."""

    def __init__(self):
        h = ctypes.c_void_p(0)
        with DEVICE_LOCK:
            st = lib().chimera_context_create(ctypes.byref(h))
        check("chimera_context_create", st)
        self.handle = h.value
        self._prev = []

    def __enter__(self):
        self._prev.append(lib().chimera_context_current())
        check("chimera_context_bind", lib().chimera_context_bind(self.handle))
        return self

    def __exit__(self, *exc):
        check("chimera_context_bind", lib().chimera_context_bind(self._prev.pop()))
        return False

    def done(self) -> bool:
        d = ctypes.c_int32(0)
        check("chimera_context_query", lib().chimera_context_query(self.handle, ctypes.byref(d)))
        return bool(d.value)

    def synchronize(self):
        with self:
            check("chimera_synchronize", lib().chimera_synchronize())

    def close(self):
        if self.handle:
            with DEVICE_LOCK:
                st = lib().chimera_context_destroy(self.handle)
            check("chimera_context_destroy", st)
            self.handle = None


class Simulation:
    """One simulation on its own context, advanced asynchronously.

    build() constructs the ChimeraRun (its arrays are allocated and its
    initial half step runs on this simulation's context).  advance(n) queues n
    more steps and returns at once; done() says whether everything queued has
    finished; result() waits and returns the run."""

    def __init__(self, build, steps, graphs=True, name="", start=0):
        self.name = name
        self.steps = int(steps)
        self.start = int(start)         # steps already done (a run resumed from a snapshot)
        self.ctx = Context()
        self.t_admit = time.perf_counter()
        with self.ctx:
            self.run = build()
            self.graphs_on = bool(graphs) and not self.run.graph_blockers()
            if self.graphs_on:
                self.run.enable_device_frame()
        self.graphs = {}
        self.istep = self.start
        self.t_done = None

    @property
    def particles(self):
        return sum(sp.np for sp in self.run.Particles if "Still" not in sp.Args["Features"])

    def _step(self, i):
        L = lib()
        if not self.graphs_on or i <= self.start + 2:
            self.run.make_step(i)
            return
        parity = i % 2
        g = self.graphs.get(parity)
        if g is None:
            # no other thread may touch the device while the step is recorded
            with DEVICE_LOCK:
                check("chimera_capture_begin", L.chimera_capture_begin())
                try:
                    self.run.make_step(i)
                except Exception:
                    L.chimera_capture_abort()
                    raise
                h = ctypes.c_void_p(0)
                check("chimera_capture_end", L.chimera_capture_end(ctypes.byref(h)))
            g = self.graphs[parity] = h.value
        else:
            self.run.host_step(i)
        # like every other library call, under the device lock: no builder
        # thread allocates, copies or captures while the graph is submitted
        with DEVICE_LOCK:
            st = L.chimera_graph_launch(g)
        check("chimera_graph_launch", st)

    def advance(self, n=1):
        """Queue up to n more steps (fewer at the end of the run); returns how many."""
        k = min(int(n), self.steps - self.istep)
        if k <= 0:
            return 0
        with self.ctx:
            for _ in range(k):
                self._step(self.istep + 1)
                self.istep += 1
        return k

    @property
    def finished_issuing(self):
        return self.istep >= self.steps

    def done(self) -> bool:
        if self.t_done is not None:
            return True
        if self.finished_issuing and self.ctx.done():
            self.t_done = time.perf_counter()
            return True
        return False

    def result(self):
        self.ctx.synchronize()
        if self.t_done is None and self.finished_issuing:
            self.t_done = time.perf_counter()
        return self.run

    def close(self):
        """Free the graphs, the simulation's arrays and its context."""
        L = lib()
        self.ctx.synchronize()
        for g in self.graphs.values():
            with DEVICE_LOCK:
                st = L.chimera_graph_destroy(g)
            check("chimera_graph_destroy", st)
        self.graphs = {}
        with self.ctx:
            self.run = None      # arrays are freed while the context is current
        self.ctx.close()


def run_batch(sims, chunk=1):
    """Advance every simulation to its end, issuing `chunk` steps of each in
    turn so that the simulations' work overlaps on the device; returns when
    all have finished."""
    active = list(sims)
    while any(not s.finished_issuing for s in active):
        for s in active:
            s.advance(chunk)
    for s in active:
        s.result()
    return sims


def free_device_bytes():
    f, t = ctypes.c_int64(0), ctypes.c_int64(0)
    check("chimera_mem_info", lib().chimera_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value
