"""Scheduling logic of the serving loop, driven by a host-only backend: admission by budget and
slots, rejection of requests that can never fit, continuous batching (late arrivals join while
others run), slot reuse, eviction on a health flag, and release of every reservation."""

import threading
import time

import numpy as np
import pytest

from fdtdx_amd import serving as S


class FakeClock:
    def __init__(self):
        self.t = 0.0
        self._lock = threading.Lock()

    def __call__(self):
        with self._lock:
            self.t += 1e-3  # time passes while the loop polls
            return self.t


class FakeBackend:
    """One chunk costs one time unit; a request's flux trace is its rid repeated; the request
    `nan_rid` raises the non-finite flag once it passes step `nan_step`."""

    def __init__(self, clock, nan_rid=None, nan_step=5):
        self.clock = clock
        self.nan_rid, self.nan_step = nan_rid, nan_step
        self.attached = {}
        self.released = []
        self.max_attached = 0
        self.advance_calls = 0

    def estimate(self, req):
        return req.cells

    def prepare(self, req):
        return {"req": req, "n": 0, "flags": 0}

    def attach(self, slot, job):
        assert slot not in self.attached
        self.attached[slot] = job
        self.max_attached = max(self.max_attached, len(self.attached))

    def detach(self, slot, job):
        assert self.attached.pop(slot) is job

    def advance(self, steps):
        self.advance_calls += 1
        time.sleep(2e-4)  # the host waits on the device here, letting preparation threads run
        self.clock.t += 1.0
        for job in self.attached.values():
            job["n"] = min(job["req"].steps, job["n"] + steps)
            if job["req"].rid == self.nan_rid and job["n"] >= self.nan_step:
                job["flags"] |= 1
        return object()

    def step_index(self, slot, job):
        return job["n"]

    def is_complete(self, token):
        return True

    def wait(self, token):
        pass

    def flags_of(self, slot, job):
        return job["flags"]

    def finish(self, job):
        r = job["req"]
        return np.full((r.steps, 1), float(r.rid), np.float32), None

    def start_reverse(self, job, token):
        return job

    def reverse_done(self, h):
        return True

    def reverse_result(self, job):
        r = job["req"]
        return np.full((r.steps, 1), float(r.rid), np.float32), np.full((2, 2, 2), -float(r.rid), np.float32)

    def release(self, job):
        self.released.append(job["req"].rid)

    # requests running alone (Server(solo_cells=...)): done after two polls, never in the batch
    def start_solo(self, job):
        self.solo_started = getattr(self, "solo_started", []) + [job["req"].rid]
        self.solo_running = getattr(self, "solo_running", 0) + 1
        self.max_solo = max(getattr(self, "max_solo", 0), self.solo_running)
        job["polls"] = 0
        return job

    def solo_done(self, job):
        job["polls"] += 1
        return job["polls"] >= 2

    def solo_result(self, job):
        self.solo_running -= 1
        r = job["req"]
        grad = np.full((2, 2, 2), -float(r.rid), np.float32) if r.gradient else None
        return np.full((r.steps, 1), float(r.rid), np.float32), grad


def _req(rid, n=8, steps=40, arrival=0.0, grad=False):
    return S.Request(rid, (n, n, n), steps, rid, "dipole", grad, arrival, 2)


def _run(reqs, capacity=10 ** 9, max_slots=64, chunk=8, policy="sjf", max_active_cells=None, prep_workers=2,
         solo_cells=None, solo_max=2, **kw):
    clock = FakeClock()
    be = FakeBackend(clock, **kw)
    budget = S.HbmBudget(capacity)
    srv = S.Server(be, budget, max_slots=max_slots, chunk=chunk, clock=clock, idle_sleep=0.0, policy=policy,
                   max_active_cells=max_active_cells, prep_workers=prep_workers, solo_cells=solo_cells,
                   solo_max=solo_max)
    res, stats = srv.run(reqs)
    return res, stats, be, budget


def test_every_request_completes_with_its_own_result():
    reqs = [_req(i, n=4 + i % 3, steps=10 + 7 * i, grad=(i % 3 == 0)) for i in range(20)]
    res, stats, be, budget = _run(reqs)
    assert sorted(res) == list(range(20))
    for r in reqs:
        x = res[r.rid]
        assert x.status == "ok"
        assert x.flux.shape == (r.steps, 1) and np.all(x.flux == r.rid)
        assert (x.grad is not None) == r.gradient
    assert sorted(be.released) == list(range(20))
    assert budget.used == 0 and not be.attached


def test_oversized_request_is_rejected_and_others_run():
    reqs = [_req(0, n=4), _req(1, n=20), _req(2, n=4)]
    res, stats, be, budget = _run(reqs, capacity=1000)
    assert res[1].status == "rejected" and "budget" in res[1].reason
    assert res[0].status == res[2].status == "ok"
    assert stats.rejected == 1


def test_budget_bounds_concurrency():
    reqs = [_req(i, n=10, steps=30) for i in range(6)]  # 1000 cells each
    res, stats, be, budget = _run(reqs, capacity=2500)
    assert all(x.status == "ok" for x in res.values())
    assert be.max_attached <= 2
    assert stats.peak_bytes <= 2500


def test_slots_are_reused():
    reqs = [_req(i, steps=16 + 8 * (i % 2)) for i in range(7)]
    res, stats, be, budget = _run(reqs, max_slots=2)
    assert all(x.status == "ok" for x in res.values())
    assert be.max_attached <= 2 and stats.peak_slots <= 2


def test_late_arrival_joins_a_running_batch():
    long_req = _req(0, steps=800, arrival=0.0)
    short = _req(1, steps=16, arrival=5.0)
    res, stats, be, budget = _run([long_req, short], chunk=8)
    assert res[1].done < res[0].done          # the short request did not wait for the long one
    assert res[1].arrival >= 5.0
    assert be.max_attached == 2


def test_flagged_slot_is_evicted_and_released():
    reqs = [_req(i, steps=40) for i in range(4)]
    res, stats, be, budget = _run(reqs, nan_rid=2)
    assert res[2].status == "failed" and "non-finite" in res[2].reason
    assert all(res[i].status == "ok" for i in (0, 1, 3))
    assert stats.evicted == 1 and 2 in be.released and budget.used == 0


def test_request_mix_is_seeded_and_within_ranges():
    a = S.make_requests(64, rate=10.0, seed=3)
    b = S.make_requests(64, rate=10.0, seed=3)
    assert [(r.shape, r.steps, r.seed, r.source, r.gradient, r.arrival) for r in a] == \
           [(r.shape, r.steps, r.seed, r.source, r.gradient, r.arrival) for r in b]
    assert all(r.shape in S.SIZES and 500 <= r.steps <= 4000 for r in a)
    arr = [r.arrival for r in a]
    assert arr == sorted(arr) and arr[0] > 0
    assert 0.05 < np.mean([r.gradient for r in a]) < 0.6


def test_fifo_latency_of_a_serial_server():
    reqs = [_req(0, arrival=0.0), _req(1, arrival=0.5), _req(2, arrival=10.0)]
    lat = S.fifo_latencies(reqs, {0: 2.0, 1: 2.0, 2: 1.0})
    assert lat == pytest.approx({0: 2.0, 1: 3.5, 2: 1.0})


def test_estimate_covers_the_gradient_record():
    f = S.Request(0, (64, 64, 64), 1000, 0, "dipole", False)
    g = S.Request(0, (64, 64, 64), 1000, 0, "dipole", True)
    faces = 2 * 3 * 64 * 64
    assert S.estimate_bytes(g) - S.estimate_bytes(f) >= 4 * 1000 * 6 * faces


def test_shortest_job_first_order_with_one_slot():
    reqs = [_req(0, steps=300), _req(1, steps=40), _req(2, steps=120), _req(3, steps=80, grad=True)]
    res, stats, be, budget = _run(reqs, max_slots=1, prep_workers=1)  # one worker: deterministic order
    order = sorted(res, key=lambda rid: res[rid].done)
    work = {r.rid: S.estimate_work(r) for r in reqs}
    assert order == sorted(work, key=work.get)  # least device work first
    res_f, _, _, _ = _run(reqs, max_slots=1, policy="fifo", prep_workers=1)
    assert sorted(res_f, key=lambda rid: res_f[rid].done) == [0, 1, 2, 3]


def test_active_cell_cap_bounds_concurrency():
    reqs = [_req(i, n=10, steps=50) for i in range(6)]  # 1000 cells each
    res, stats, be, budget = _run(reqs, max_active_cells=2500)
    assert all(x.status == "ok" for x in res.values())
    assert be.max_attached <= 2


def test_large_requests_run_alone_beside_the_batch():
    """solo_cells: requests of at least that many cells never enter the batch, at most solo_max run
    alone at a time, and every request still returns its own result and releases its budget."""
    reqs = [_req(i, n=(12 if i % 4 == 0 else 5), steps=10 + 3 * i, grad=(i % 3 == 0)) for i in range(16)]
    res, stats, be, budget = _run(reqs, solo_cells=12 ** 3, solo_max=2)
    big = sorted(r.rid for r in reqs if r.cells >= 12 ** 3)
    assert sorted(be.solo_started) == big
    assert be.max_solo <= 2
    for r in reqs:
        x = res[r.rid]
        assert x.status == "ok" and np.all(x.flux == r.rid)
        assert (x.grad is not None) == r.gradient
    assert sorted(be.released) == list(range(16))
    assert budget.used == 0 and not be.attached
    solo_events = {rid for _, rid, stage in stats.events if stage == "solo"}
    attach_events = {rid for _, rid, stage in stats.events if stage == "attach"}
    assert solo_events == set(big) and not (attach_events & set(big))
