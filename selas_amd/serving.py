"""Request loop for VCM render jobs on one GPU: admission by HBM footprint, resident scenes and
continuous batching.

Selas renders one job at a time from its application loop (Source/Applications/Selas/Source/
Main.cpp, VCM::GenerateImage at VCM.cpp:669-760); it has no request loop. This one serves a stream
of render requests (a scene id, an image size, a number of iterations, a seed):

* Scenes are kept resident on the device (tree, triangles, shading records, textures) and reused by
  every request on them; a scene is built on the first request that needs it, and evicted, least
  recently used first, only when memory is short and no admitted job uses it.
* A request is admitted, in arrival order, when an admission slot is free and its footprint (its
  framebuffer, plus its scene when that is not resident) fits the HBM budget beside everything
  already admitted or resident; the head of the queue waits for memory to free up while jobs are in
  flight, and is rejected only when it cannot fit even on an idle device.
* Admitted jobs share the engine's wavefronts (sgpu_vcm_server_step): work of requests that arrive
  while others are on the GPU is packed into the same launches, so they overlap on the device.

The engine is duck-typed (GpuEngine below drives the C ABI; tests drive the loop with a simulated
engine and clock): job_bytes(r), scene_bytes(sid), resident_scenes(), load_scene(sid),
evict_scene(sid), submit(r) -> handle, step() -> handles retired, collect(handle) -> image.
"""
from __future__ import annotations

import math
import time
from collections import OrderedDict, deque
from dataclasses import dataclass

import numpy as np


@dataclass
class Request:
    id: int
    arrival: float            # seconds after the start of the run
    scene: int                # scene id
    width: int
    height: int
    iterations: int
    seed: int
    max_path_length: int = 10
    admitted: float = math.nan
    finished: float = math.nan
    status: str = "pending"   # "running", "done" or "rejected"
    image: object = None

    @property
    def samples(self) -> int:
        return self.width * self.height * self.iterations

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def poisson_requests(seed: int, rate: float, jobs) -> list:
    """Requests for `jobs` (dicts with scene, width, height, iterations, seed, max_path_length, as
    workloads.job_mix makes them) arriving as a Poisson process of `rate` requests per second."""
    rng = np.random.default_rng((seed, int(round(rate * 1000))))
    t = np.cumsum(rng.exponential(1.0 / rate, len(jobs)))
    return [Request(k, float(t[k]), int(j["scene"]), int(j["width"]), int(j["height"]), int(j["iterations"]),
                    int(j["seed"]), int(j.get("max_path_length", 10))) for k, j in enumerate(jobs)]


class RequestLoop:
    """Admits and runs requests against an engine in real time (clock and sleep are injectable).

    hbm_budget bounds the bytes of resident scenes plus admitted framebuffers (the engine's fixed
    wavefront buffers are outside it); max_slots bounds the jobs admitted at once."""

    def __init__(self, engine, hbm_budget: int, max_slots: int = 64, keep_images: bool = False,
                 clock=time.perf_counter, sleep=time.sleep):
        if max_slots < 1:
            raise ValueError("max_slots must be at least 1")
        self.engine = engine
        self.budget = int(hbm_budget)
        self.max_slots = max_slots
        self.keep_images = keep_images
        self.clock = clock
        self.sleep = sleep
        self.lru = OrderedDict()  # resident scene id -> bytes charged, least recently used first
        for sid in engine.resident_scenes():
            self.lru[sid] = engine.scene_bytes(sid)
        self.used = sum(self.lru.values())
        self.peak_used = self.used
        self.steps = 0
        self.busy = 0.0
        self.makespan = 0.0
        self.evictions = 0
        self.peak_in_flight = 0

    def _make_room(self, need: int, keep: int, scene_jobs: dict):
        for sid in list(self.lru):
            if self.used + need <= self.budget:
                return
            if sid != keep and scene_jobs.get(sid, 0) == 0:
                self.used -= self.lru.pop(sid)
                self.engine.evict_scene(sid)
                self.evictions += 1

    def run(self, requests) -> list:
        e = self.engine
        reqs = sorted(requests, key=lambda r: (r.arrival, r.id))
        t0 = self.clock()
        now = lambda: self.clock() - t0
        nxt = 0
        queue = deque()
        flight = {}      # engine handle -> request
        scene_jobs = {}  # scene id -> admitted jobs on it
        while nxt < len(reqs) or queue or flight:
            t = now()
            while nxt < len(reqs) and reqs[nxt].arrival <= t:
                queue.append(reqs[nxt])
                nxt += 1
            # admission, in arrival order
            while queue and len(flight) < self.max_slots:
                r = queue[0]
                jb = e.job_bytes(r)
                need = jb + (0 if r.scene in self.lru else e.scene_bytes(r.scene))
                if self.used + need > self.budget:
                    self._make_room(need, r.scene, scene_jobs)
                if self.used + need > self.budget:
                    if flight:
                        break  # wait for admitted jobs to free memory
                    queue.popleft()
                    r.status = "rejected"
                    continue
                queue.popleft()
                if r.scene not in self.lru:
                    e.load_scene(r.scene)
                    self.lru[r.scene] = e.scene_bytes(r.scene)
                    self.used += self.lru[r.scene]
                self.lru.move_to_end(r.scene)
                self.used += jb
                self.peak_used = max(self.peak_used, self.used)
                h = e.submit(r)
                r.admitted = now()
                r.status = "running"
                flight[h] = r
                scene_jobs[r.scene] = scene_jobs.get(r.scene, 0) + 1
                self.peak_in_flight = max(self.peak_in_flight, len(flight))
            if flight:
                t1 = self.clock()
                done = e.step()
                self.busy += self.clock() - t1
                self.steps += 1
                for h in done:
                    r = flight.pop(h)
                    img = e.collect(h)
                    r.finished = now()
                    r.status = "done"
                    if self.keep_images:
                        r.image = img
                    self.used -= e.job_bytes(r)
                    scene_jobs[r.scene] -= 1
            elif nxt < len(reqs):
                self.sleep(max(0.0, reqs[nxt].arrival - now()))
        self.makespan = now()
        return reqs


def summarize(reqs, makespan: float) -> dict:
    """Latency percentiles (arrival to result on the host) and throughput of a finished run."""
    done = [r for r in reqs if r.status == "done"]
    lat = np.array([r.latency for r in done]) if done else np.zeros(1)
    wait = np.array([r.admitted - r.arrival for r in done]) if done else np.zeros(1)
    samples = sum(r.samples for r in done)
    last_arrival = max((r.arrival for r in reqs), default=0.0)
    return {
        "jobs": len(done),
        "rejected": sum(r.status == "rejected" for r in reqs),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3),
        "max_ms": float(lat.max() * 1e3),
        "wait_p50_ms": float(np.percentile(wait, 50) * 1e3),
        "wait_p99_ms": float(np.percentile(wait, 99) * 1e3),
        "samples": int(samples),
        "makespan_s": float(makespan),
        "samples_per_s": float(samples / makespan) if makespan > 0 else 0.0,
        "offered_samples_per_s": float(samples / last_arrival) if last_arrival > 0 else 0.0,
    }


class GpuEngine:
    """The loop's engine on the GPU: a VcmServer (sgpu_vcm_server_*), resident scenes built by
    `make_scene(sid)` (a render-scene dict, as workloads.vcm_scene makes them) and one device
    framebuffer per admitted job. `scene_tris(sid)` sizes a scene before it is built.

    plain=True gives the plain one-job-at-a-time schedule to compare with (use it with
    max_slots=1): every iteration of a job is a wavefront of its own."""

    def __init__(self, ctx, make_scene, scene_tris=None, wavefront_paths: int = 1 << 20, max_path_length: int = 10,
                 plain: bool = False, scenes: dict | None = None):
        from . import native as N
        self.N = N
        self.ctx = ctx
        self.make_scene = make_scene
        self.scene_tris = scene_tris
        flags = N.VCM_ONE_SLICE_PER_WAVEFRONT if plain else N.VCM_SHORTEST_FIRST
        self.server = N.VcmServer(ctx, wavefront_paths, max_path_length, flags)
        self.fixed_bytes = N.VcmServer.footprint(ctx, wavefront_paths, max_path_length, flags)
        # sid -> (Scene, camera, merge radius, radius alpha); `scenes` shares resident scenes
        # between engines (the caller then releases them)
        self.own_scenes = scenes is None
        self.scenes = {} if scenes is None else scenes
        self.known_bytes = {}  # sid -> device bytes measured when it was last resident
        self.jobs = {}  # handle -> (framebuffer, request)

    def close(self):
        for h in list(self.jobs):
            self.jobs.pop(h)[0].free()
        self.server.close()
        if self.own_scenes:
            for sid in list(self.scenes):
                self.evict_scene(sid)

    def resident_scenes(self):
        return list(self.scenes)

    def load_scene(self, sid):
        if sid in self.scenes:
            return
        spec = self.make_scene(sid)
        sc = self.N.build_render_scene(self.ctx, spec)
        self.scenes[sid] = (sc, spec["camera"], float(spec["merge_radius"]), float(spec.get("radius_alpha", 0.75)))

    def evict_scene(self, sid):
        sc = self.scenes.pop(sid)[0]
        self.known_bytes[sid] = sc.device_bytes()
        sc.release()

    def scene_bytes(self, sid) -> int:
        if sid in self.scenes:
            return self.scenes[sid][0].device_bytes()
        # built before: what it held then (a scene is built the same way every time)
        if sid in self.known_bytes:
            return self.known_bytes[sid]
        # never built: two 64-byte nodes and one 48-byte record per triangle, one 16 x 16 RGB8
        # texture and 32 bytes of adjacency per quad
        t = float(self.scene_tris(sid)) if self.scene_tris else 1e6
        return int(t * (2 * 64 + 48) + t / 2 * (768 + 32))

    def job_bytes(self, r: Request) -> int:
        return 12 * r.width * r.height

    def submit(self, r: Request):
        sc, cam, radius, alpha = self.scenes[r.scene]
        fb = self.ctx.empty(12 * r.width * r.height)
        job = {"camera": cam, "width": r.width, "height": r.height, "iterations": r.iterations,
               "max_path_length": r.max_path_length, "base_radius": radius, "radius_alpha": alpha, "seed": r.seed}
        try:
            h = self.server.submit(sc, job, fb)
        except Exception:
            fb.free()
            raise
        self.jobs[h] = (fb, r)
        return h

    def step(self):
        done, _ = self.server.step()
        return done

    def collect(self, h):
        fb, r = self.jobs.pop(h)
        img = fb.download(np.float32, 3 * r.width * r.height).reshape(r.height, r.width, 3)
        fb.free()
        return img
