"""Failure handling of the request loop (chimera_amd.serve.Server), forced from
the host on the CPU: simulations are stand-ins with the interface the loop
drives (advance / done / close, a context), so every path runs without a GPU.

* A build that fails (an allocation failure during admission) fails that
  request only: the others finish with the results they give alone.
* A request cancelled mid-flight frees its slot; the others are unchanged.
* A request whose result is not finite (particles that went NaN) is flagged
  as failed, not returned as a result.
"""
import time

import numpy as np
import pytest

from chimera_amd import serve as S


class FakeCtx:
    def done(self):
        return True

    def synchronize(self):
        pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


class FakeSim:
    """Advances a deterministic state: value_{i+1} = value_i * 1.5 + i."""

    def __init__(self, steps, seed, poison_at=None):
        self.steps, self.istep, self.particles = steps, 0, 100
        self.value = float(seed)
        self.poison_at = poison_at
        self.ctx = FakeCtx()
        self.t_done = None
        self.closed = False

    @property
    def finished_issuing(self):
        return self.istep >= self.steps

    def advance(self, n=1):
        assert not self.closed
        k = min(n, self.steps - self.istep)
        for _ in range(k):
            self.value = self.value * 1.5 + self.istep
            self.istep += 1
            if self.poison_at is not None and self.istep >= self.poison_at:
                self.value = float("nan")
        return k

    def done(self):
        if self.finished_issuing and self.t_done is None:
            self.t_done = time.perf_counter()
        return self.finished_issuing

    def close(self):
        self.closed = True


class FakeRequest:
    def __init__(self, name, seed, steps, t_arrival=0.0, fail=None, poison_at=None, nbytes=1 << 20):
        self.name, self.seed, self.steps, self.t_arrival = name, seed, steps, t_arrival
        self.bytes = nbytes
        self.fail, self.poison_at = fail, poison_at
        self.sim = None
        self.t_admit = self.t_built = self.t_done = None
        self.particle_steps = 0
        self.result = None
        self.snap, self.istep0, self.evictions = None, 0, 0
        self.status, self.error = "pending", None
        self.cancel_requested = False

    def build(self, graphs=True):
        if self.fail is not None:
            raise self.fail
        return FakeSim(self.steps, self.seed, self.poison_at)


class FakeServer(S.Server):
    def result_of(self, sim):
        return [np.full((3, 3), sim.value)]


def alone(r):
    s = FakeSim(r.steps, r.seed)
    s.advance(r.steps)
    return s.value


def serve(reqs, **kw):
    server = FakeServer(slots=2, reserve=0, chunk=3, free_bytes=lambda: 1 << 40, **kw)
    try:
        S.serve(reqs, server)
    finally:
        server.close()
    return server


@pytest.mark.parametrize("threads", [True, False])
def test_build_failure_fails_only_that_request(threads):
    reqs = [FakeRequest("a", 1, 20), FakeRequest("b", 2, 15, t_arrival=0.001,
                                                 fail=MemoryError("hipMalloc of 8 GiB failed")),
            FakeRequest("c", 3, 25, t_arrival=0.002)]
    server = serve(reqs, builder_thread=threads)
    assert [r.name for r in server.failed] == ["b"]
    assert reqs[1].status == "failed" and "hipMalloc" in reqs[1].error
    assert sorted(r.name for r in server.finished) == ["a", "c"]
    for r in (reqs[0], reqs[2]):
        assert r.status == "done"
        assert np.all(r.result[0] == alone(r)), r.name
    assert server.idle and server.building == 0 and server.reserved == 0


def test_cancel_mid_flight_frees_the_slot():
    reqs = [FakeRequest("a", 1, 40), FakeRequest("b", 2, 40), FakeRequest("c", 3, 30)]
    server = FakeServer(slots=2, reserve=0, chunk=2, free_bytes=lambda: 1 << 40, builder_thread=False)
    for r in reqs:
        server.submit(r, 0.0)
    server.poll(0.0)                    # a and b admitted and built; c waits for a slot
    server.poll(0.0)
    assert [r.name for r in server.active] == ["a", "b"] and [r.name for r in server.pending] == ["c"]
    sim_b = reqs[1].sim
    assert server.cancel("b")
    assert sim_b.closed and reqs[1].status == "cancelled" and reqs[1].sim is None
    assert not server.cancel("nobody")
    while not server.idle:
        server.poll(0.0)
    assert [r.name for r in server.failed] == ["b"]
    assert sorted(r.name for r in server.finished) == ["a", "c"]      # c got b's slot
    for r in (reqs[0], reqs[2]):
        assert np.all(r.result[0] == alone(r)), r.name


def test_cancel_pending_request():
    reqs = [FakeRequest("a", 1, 10), FakeRequest("b", 2, 10)]
    server = FakeServer(slots=1, reserve=0, chunk=2, free_bytes=lambda: 1 << 40, builder_thread=False)
    for r in reqs:
        server.submit(r, 0.0)
    server.poll(0.0)
    assert server.cancel("b") and reqs[1].status == "cancelled"
    while not server.idle:
        server.poll(0.0)
    assert [r.name for r in server.finished] == ["a"]


def test_non_finite_result_is_flagged():
    reqs = [FakeRequest("a", 1, 12), FakeRequest("nan", 2, 12, poison_at=5), FakeRequest("c", 3, 12)]
    server = serve(reqs, builder_thread=True)
    assert [r.name for r in server.failed] == ["nan"]
    assert reqs[1].status == "failed" and "non-finite" in reqs[1].error
    assert sorted(r.name for r in server.finished) == ["a", "c"]


def test_admission_waits_for_memory():
    """Requests that do not fit the free memory wait in arrival order; none is
    built before memory frees up."""
    total = {"bytes": 3 << 20}
    reqs = [FakeRequest(n, k, 6, nbytes=2 << 20) for k, n in enumerate("abc")]
    # device memory: what the built simulations hold is gone from the free bytes
    server = FakeServer(slots=8, reserve=0, chunk=2, builder_thread=False,
                        free_bytes=lambda: total["bytes"] - sum(r.bytes for r in server.active))
    for r in reqs:
        server.submit(r, 0.0)
    server.poll(0.0)
    server.poll(0.0)
    assert [r.name for r in server.active] == ["a"] and [r.name for r in server.pending] == ["b", "c"]
    total["bytes"] = 1 << 40
    while not server.idle:
        server.poll(0.0)
    assert [r.name for r in server.finished] == ["a", "b", "c"]
