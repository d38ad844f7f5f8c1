"""Serving-loop scheduling on the CPU with a stand-in backend: budget
admission, batching of one protocol's requests, eviction, and that every
request completes once."""
import threading
import time

import numpy as np
import pytest

from monalisa_amd.serve import ReconServer, percentiles


class FakeBackend:
    """Plans of protocol p and batch B cost p['plan'] * B bytes; a batch
    'runs' for p['t'] seconds and returns each request's y doubled."""

    def __init__(self):
        self.live = {}
        self.lock = threading.Lock()
        self.max_live = 0
        self.concurrent = 0
        self.max_concurrent = 0

    def batch_capacity(self, proto):
        return proto.get("cap", 99)

    def plan_bytes_estimate(self, proto, B):
        return proto["plan"] * B

    def request_bytes(self, proto):
        return proto["req"]

    def new_stream(self):
        return None

    def make_plan(self, proto, B, stream):
        p = object()
        with self.lock:
            self.live[id(p)] = proto["plan"] * B
            self.max_live = max(self.max_live, sum(self.live.values()))
        return p

    def plan_bytes(self, plan):
        return self.live[id(plan)]

    def run(self, plan, proto, reqs, stream):
        with self.lock:
            self.concurrent += 1
            self.max_concurrent = max(self.max_concurrent, self.concurrent)
        time.sleep(proto["t"])
        for r in reqs:
            r.x = 2 * r.y
        with self.lock:
            self.concurrent -= 1

    def release(self, plan):
        with self.lock:
            self.live.pop(id(plan), None)


def _server(budget, max_batch=4, workers=3):
    be = FakeBackend()
    s = ReconServer(be, budget, max_batch=max_batch, workers=workers)
    s.register("small", {"plan": 100, "req": 1, "t": 0.01})
    s.register("big", {"plan": 1000, "req": 10, "t": 0.03})
    return s, be


def test_every_request_completes_once_with_its_own_result():
    s, be = _server(budget=10_000)
    s.start()
    reqs = [s.submit("small" if i % 3 else "big", np.full(3, i)) for i in range(40)]
    assert s.drain(timeout=30)
    s.stop()
    assert len(s.completed) == 40
    assert sorted(r.rid for r in s.completed) == list(range(40))
    for i, r in enumerate(reqs):
        assert r.done.is_set() and r.error is None
        np.testing.assert_array_equal(r.x, np.full(3, 2 * i))
        assert r.t_submit <= r.t_start <= r.t_done


def test_batch_capacity_per_protocol():
    s, be = _server(budget=100_000, max_batch=4, workers=1)
    s.register("solo", {"plan": 100, "req": 1, "t": 0.005, "cap": 1})
    reqs = [s.submit("solo", np.zeros(1)) for _ in range(3)]
    s.start()
    assert s.drain(timeout=30)
    s.stop()
    assert [b for _, b in s.batches] == [1, 1, 1]
    assert s.capacity("solo") == 1 and s.capacity("small") == 4


def test_requests_of_one_protocol_are_batched_and_batches_overlap():
    s, be = _server(budget=100_000, max_batch=4, workers=3)
    # queue everything before the dispatcher starts: batches of 4 form
    reqs = [s.submit("small", np.zeros(1)) for _ in range(12)]
    s.start()
    assert s.drain(timeout=30)
    s.stop()
    assert [b for _, b in s.batches] == [4, 4, 4]
    assert be.max_concurrent >= 2
    assert all(r.batch == 4 for r in reqs)


def test_budget_is_never_exceeded_and_idle_plans_are_evicted():
    # one big plan of batch 4 (4000 + 40) fits, two do not
    s, be = _server(budget=5_000, max_batch=4, workers=3)
    s.start()
    for i in range(30):
        s.submit("big" if i % 2 else "small", np.zeros(1))
        time.sleep(0.002)
    assert s.drain(timeout=60)
    s.stop()
    assert s.peak_bytes <= 5_000
    assert be.max_live <= 5_000
    assert len(s.completed) == 30


def test_oldest_request_first():
    s, be = _server(budget=100_000, max_batch=1, workers=1)
    order = ["big", "small", "big", "small"]
    reqs = [s.submit(k, np.zeros(1)) for k in order]
    s.start()
    assert s.drain(timeout=30)
    s.stop()
    starts = [r.t_start for r in reqs]
    assert starts == sorted(starts)


def test_unknown_protocol_and_percentiles():
    s, _ = _server(budget=1000)
    with pytest.raises(KeyError):
        s.submit("nope", np.zeros(1))
    p = percentiles([1, 2, 3, 4, 100])
    assert p["p50"] == 3.0 and 4.0 < p["p99"] <= 100.0


def test_request_larger_than_budget_fails_instead_of_blocking():
    s, _ = _server(budget=500)
    s.start()
    r_big = s.submit("big", np.zeros(1))
    r_small = s.submit("small", np.ones(1))
    assert s.drain(timeout=30)
    s.stop()
    assert isinstance(r_big.error, MemoryError)
    assert r_small.error is None and r_small.x[0] == 2


def test_batches_never_mix_cg_step_counts():
    """A batch runs one nCGD (the backend takes it from the batch), so
    requests with different nCGD for one protocol go to different batches,
    and each result is the one the request would get alone."""
    be = FakeBackend()
    seen = []
    run0 = be.run

    def run(plan, proto, reqs, stream):
        seen.append([r.nCGD for r in reqs])
        run0(plan, proto, reqs, stream)
        for r in reqs:
            r.x = r.x * r.nCGD          # the result depends on the request's own nCGD

    be.run = run
    s = ReconServer(be, 10_000, max_batch=4, workers=1)
    s.register("small", {"plan": 100, "req": 1, "t": 0.002})
    reqs = [s.submit("small", np.full(2, i), nCGD=(5 if i % 3 else 10)) for i in range(30)]
    s.start()
    assert s.drain(timeout=30)
    s.stop()
    assert all(len(set(b)) == 1 for b in seen), seen
    assert max(len(b) for b in seen) > 1
    for i, r in enumerate(reqs):
        assert r.error is None
        np.testing.assert_array_equal(r.x, np.full(2, 2 * i) * r.nCGD)
    assert sorted(r.rid for r in s.completed) == list(range(30))
