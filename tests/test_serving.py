"""The request loop's admission and batching logic (selas_amd/serving.py) against a simulated
engine on a simulated clock: every wavefront takes `dt` seconds and carries up to `capacity`
samples of the admitted jobs in admission order."""
import math

import numpy as np
import pytest

from selas_amd import serving as S
from selas_amd import workloads as W


class Clock:
    def __init__(self):
        self.t = 0.0
        self.sleeps = 0

    def clock(self):
        return self.t

    def sleep(self, dt):
        assert dt >= 0
        self.sleeps += 1
        self.t += dt


class SimEngine:
    def __init__(self, clock, scene_bytes, resident=(), capacity=1 << 20, dt=0.02, max_slots=None):
        self.c = clock
        self.sb = dict(scene_bytes)
        self.res = set(resident)
        self.capacity = capacity
        self.dt = dt
        self.max_slots = max_slots
        self.active = []  # [handle, request, samples left]
        self.next = 0
        self.loads, self.evicts = [], []

    def resident_scenes(self):
        return sorted(self.res)

    def scene_bytes(self, sid):
        return self.sb[sid]

    def job_bytes(self, r):
        return 12 * r.width * r.height

    def load_scene(self, sid):
        assert sid not in self.res
        self.res.add(sid)
        self.loads.append(sid)

    def evict_scene(self, sid):
        assert all(r.scene != sid for _, r, _ in self.active), "evicted a scene in use"
        self.res.remove(sid)
        self.evicts.append(sid)

    def submit(self, r):
        assert r.scene in self.res
        self.active.append([self.next, r, r.samples])
        if self.max_slots:
            assert len(self.active) <= self.max_slots
        self.next += 1
        return self.next - 1

    def step(self):
        self.c.t += self.dt
        cap = self.capacity
        done = []
        for a in self.active:
            take = min(cap, a[2])
            a[2] -= take
            cap -= take
            if a[2] == 0:
                done.append(a[0])
        self.active = [a for a in self.active if a[2] > 0]
        return done

    def collect(self, h):
        return h


def _req(k, t, scene=0, side=64, its=4):
    return S.Request(k, t, scene, side, side, its, k)


def test_every_request_served_and_latency_bookkeeping():
    c = Clock()
    e = SimEngine(c, {0: 1000, 1: 1000}, resident=[0, 1], max_slots=4)
    jobs = W.job_mix(3, 40, n_scenes=2)
    reqs = S.poisson_requests(1, 20.0, jobs)
    loop = S.RequestLoop(e, hbm_budget=1 << 40, max_slots=4, clock=c.clock, sleep=c.sleep)
    out = loop.run(reqs)
    assert all(r.status == "done" for r in out)
    for r in out:
        assert r.arrival <= r.admitted < r.finished
        # at least the wavefronts its own samples need
        assert r.finished - r.admitted >= e.dt * math.ceil(r.samples / e.capacity) - 1e-9
    assert loop.peak_in_flight <= 4
    st = S.summarize(out, loop.makespan)
    assert st["jobs"] == 40 and st["rejected"] == 0
    assert st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["samples"] == sum(r.samples for r in out)


def test_arrivals_during_a_job_are_batched_with_it():
    c = Clock()
    e = SimEngine(c, {0: 10}, resident=[0], capacity=64 * 64, dt=0.01)
    big = _req(0, 0.0, side=64, its=50)     # 50 wavefronts of its own
    small = _req(1, 0.105, side=8, its=1)   # arrives while it runs
    loop = S.RequestLoop(e, hbm_budget=1 << 30, max_slots=8, clock=c.clock, sleep=c.sleep)
    loop.run([big, small])
    assert small.admitted < big.finished   # admitted while the big one is in flight
    # wavefronts of 10 ms from t = 0, filled in admission order: 50 for the big job, then the small one
    assert big.finished == pytest.approx(0.50) and small.finished == pytest.approx(0.51)
    # one slot: the small one waits for the big one to finish
    c2 = Clock()
    e2 = SimEngine(c2, {0: 10}, resident=[0], capacity=64 * 64, dt=0.01)
    big2, small2 = _req(0, 0.0, side=64, its=50), _req(1, 0.105, side=8, its=1)
    S.RequestLoop(e2, hbm_budget=1 << 30, max_slots=1, clock=c2.clock, sleep=c2.sleep).run([big2, small2])
    assert small2.admitted >= big2.finished


def test_idle_loop_sleeps_until_the_next_arrival():
    c = Clock()
    e = SimEngine(c, {0: 10}, resident=[0], dt=0.001)
    reqs = [_req(0, 1.0), _req(1, 5.0)]
    loop = S.RequestLoop(e, hbm_budget=1 << 30, clock=c.clock, sleep=c.sleep)
    loop.run(reqs)
    assert reqs[0].admitted == pytest.approx(1.0) and reqs[1].admitted == pytest.approx(5.0)
    assert loop.steps == 2 and c.sleeps == 2


def test_admission_by_footprint_waits_for_memory():
    c = Clock()
    e = SimEngine(c, {0: 0}, resident=[0], capacity=100, dt=0.01)
    fb = 12 * 64 * 64
    reqs = [_req(k, 0.0, side=64, its=2) for k in range(5)]
    loop = S.RequestLoop(e, hbm_budget=2 * fb, max_slots=8, clock=c.clock, sleep=c.sleep)
    loop.run(reqs)
    assert loop.peak_in_flight == 2 and loop.peak_used <= 2 * fb
    assert all(r.status == "done" for r in reqs)
    # admitted in arrival order, never more than two at once
    adm = sorted(reqs, key=lambda r: r.admitted)
    assert [r.id for r in adm] == list(range(5))
    fin = sorted(r.finished for r in reqs)
    for k in range(2, 5):  # the k-th admission waits for the (k-2)-th completion
        assert adm[k].admitted >= fin[k - 2]


def test_scenes_loaded_on_demand_and_evicted_lru_when_idle():
    c = Clock()
    sb = {0: 1000, 1: 1000, 2: 1000}
    e = SimEngine(c, sb, resident=[], capacity=1 << 20, dt=0.01)
    fb = 12 * 8 * 8
    # room for two scenes and two framebuffers
    loop = S.RequestLoop(e, hbm_budget=2000 + 2 * fb, max_slots=4, clock=c.clock, sleep=c.sleep)
    reqs = [_req(0, 0.0, scene=0, side=8), _req(1, 0.1, scene=1, side=8), _req(2, 0.2, scene=0, side=8),
            _req(3, 0.3, scene=2, side=8), _req(4, 0.4, scene=0, side=8)]
    loop.run(reqs)
    assert all(r.status == "done" for r in reqs)
    assert e.loads == [0, 1, 2]          # scene 0 stays resident for request 2
    assert e.evicts == [1]               # least recently used idle scene makes room for scene 2
    assert loop.peak_used <= loop.budget


def test_request_that_never_fits_is_rejected():
    c = Clock()
    e = SimEngine(c, {0: 5000, 1: 10}, resident=[1])
    loop = S.RequestLoop(e, hbm_budget=4000, clock=c.clock, sleep=c.sleep)
    reqs = [_req(0, 0.0, scene=0, side=8), _req(1, 0.0, scene=1, side=8)]
    loop.run(reqs)
    assert reqs[0].status == "rejected" and reqs[1].status == "done"
    assert S.summarize(reqs, loop.makespan)["rejected"] == 1


def test_poisson_requests_rate():
    jobs = W.job_mix(1, 4000, n_scenes=16)
    reqs = S.poisson_requests(2, 50.0, jobs)
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert abs(gaps.mean() - 1 / 50.0) < 0.1 / 50.0
    assert {r.scene for r in reqs} == set(range(16))
    sides = np.array([r.width for r in reqs])
    assert sides.min() >= 64 and sides.max() <= 1024
    its = np.array([r.iterations for r in reqs])
    assert its.min() == 1 and its.max() == 64
    # log-uniform side: the median is near sqrt(64 * 1024) = 256
    assert 220 < np.median(sides) < 300
