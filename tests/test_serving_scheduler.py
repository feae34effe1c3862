"""The serving loop's scheduling on the CPU: admission against the HBM budget,
eviction, continuous batching of right-hand sides and progress when the
budget holds fewer problems than have waiting requests. The residents here
are stand-ins (a footprint proportional to N and a solve that sleeps and
returns its right-hand sides), so only the scheduler is under test."""
import threading
import time

import numpy as np
import pytest

from mxc_amd.serving import ProblemSpec, Server, latency_stats, poisson_load


class FakeResident:
    live = 0
    peak = 0
    lock = threading.Lock()

    def __init__(self, pid, spec, device, max_rhs):
        self.pid, self.spec, self.max_rhs = pid, spec, max_rhs
        self.footprint = 1000 * spec.n
        self.last_used = time.perf_counter()
        time.sleep(0.01)
        with FakeResident.lock:
            FakeResident.live += self.footprint
            FakeResident.peak = max(FakeResident.peak, FakeResident.live)

    def solve(self, B, tol):
        assert B.shape[1] <= self.max_rhs
        time.sleep(0.002 * B.shape[1])
        return B.copy(), [0.0] * B.shape[1]

    def close(self):
        with FakeResident.lock:
            FakeResident.live -= self.footprint


SIZES = [16384, 32768, 65536, 16384, 32768, 65536]


def make(budget, slots, max_rhs=8):
    FakeResident.live = FakeResident.peak = 0
    srv = Server(hbm_budget=budget, slots=slots, max_rhs=max_rhs, resident_factory=FakeResident)
    for pid, n in enumerate(SIZES):
        srv.register(pid, ProblemSpec(n=n, seed=pid))
    return srv


@pytest.mark.timeout(60)
def test_every_request_served_under_a_tight_budget():
    # room for about two of the largest problems: every round of requests
    # needs evictions, and residents with waiting requests must still be served
    srv = make(budget=int(2.2 * 1000 * 65536), slots=2)
    try:
        rng = np.random.default_rng(0)
        reqs = []
        for rnd in range(3):
            for pid, n in enumerate(SIZES):
                reqs.append(srv.submit(pid, rng.uniform(-1, 1, size=(n, 2))))
            for r in reqs:
                r.wait(20)
            assert srv.resident_bytes() <= srv.budget
        assert srv.evictions > 0
        assert FakeResident.peak <= srv.budget
        for r in reqs:
            assert np.array_equal(r.X, r.B)
    finally:
        srv.close()
    assert FakeResident.live == 0


@pytest.mark.timeout(60)
def test_requests_join_calls_and_keep_their_columns():
    srv = make(budget=1 << 40, slots=2, max_rhs=16)
    try:
        reqs = poisson_load(srv, list(range(len(SIZES))), SIZES, rate=400.0, count=60, seed=1,
                            max_rhs_per_request=6, timeout=20)
        st = latency_stats(reqs, srv.calls)
        assert st["requests"] == 60
        assert srv.evictions == 0
        assert any(c[2] > 1 for c in srv.calls)  # some calls carried several requests
        assert all(c[1] <= 16 for c in srv.calls)
        for r in reqs:
            assert np.array_equal(r.X, r.B)
            assert r.t_submit <= r.t_start <= r.t_done
    finally:
        srv.close()


@pytest.mark.timeout(60)
def test_problem_larger_than_the_budget_is_refused():
    srv = make(budget=1000 * 20000, slots=1)
    try:
        ok = srv.submit(0, np.ones((SIZES[0], 1)))
        assert np.array_equal(ok.wait(20), np.ones((SIZES[0], 1)))
        bad = srv.submit(2, np.ones((SIZES[2], 1)))
        with pytest.raises(MemoryError):
            bad.wait(20)
    finally:
        srv.close()


def test_bad_requests_raise():
    srv = make(budget=1 << 40, slots=1, max_rhs=4)
    try:
        with pytest.raises(ValueError):
            srv.submit(0, np.ones((SIZES[0], 5)))
        with pytest.raises(ValueError):
            srv.submit(0, np.ones((SIZES[0] + 1, 1)))
        with pytest.raises(KeyError):
            srv.submit(99, np.ones((10, 1)))
    finally:
        srv.close()
