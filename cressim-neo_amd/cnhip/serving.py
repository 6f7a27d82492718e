"""Continuous-batching step server for many independent scenes resident in HBM.

Sessions are admitted against a device-memory budget and placed, balanced by particle count, in
a few shard pools.  Each shard is one context (one HIP stream) holding its sessions as a scene
pool.  The loop takes step requests as they arrive; whenever a shard is idle it steps, in one
launch per pass (``cn_step_scenes``, graph-captured), every session of the shard that has a pending
request, at most one request per session per tick so a session's steps stay in arrival order.
Requests that arrive while a tick runs wait for the next one; ticks of different shards run on
different streams and overlap on the device.  A request's latency runs from its arrival to the
host observing the completion of the tick that served it.

The reference has no serving loop; its RL harness steps whole worlds from Python
(examples/python/throughput_benchmark.py).  The engine calls are the C ABI's.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass

import numpy as np

from . import Scene, Solver

# Device bytes per particle of a resident scene (state double-buffered in slot order, broadphase,
# fluid records, neighbor lists at 64 slots and candidate lists at 96, halo descriptors), per
# particle of a scene with solid particles (contact candidates and records, 32 per particle) and
# per soft constraint.
BYTES_PER_PARTICLE = 840
BYTES_PER_CONTACT_PARTICLE = 664
BYTES_PER_CONSTRAINT = 80


def estimate_bytes(sc: Scene) -> int:
    cons = sum(len(x) for x in (sc.edges, sc.bends, sc.dihedrals, sc.tets) if x is not None)
    contact = BYTES_PER_CONTACT_PARTICLE if bool((np.asarray(sc.kinds) == 0).any()) else 0
    return int(sc.n * (BYTES_PER_PARTICLE + contact) + cons * BYTES_PER_CONSTRAINT)


def admit(scenes, budget_bytes: int):
    """Greedy admission in session order: (admitted session ids, deferred session ids, bytes used)."""
    admitted, deferred, used = [], [], 0
    for k, sc in enumerate(scenes):
        b = estimate_bytes(sc)
        if used + b <= budget_bytes:
            admitted.append(k)
            used += b
        else:
            deferred.append(k)
    return admitted, deferred, used


def shard_sessions(sizes, ids, shards: int):
    """Longest-first assignment of sessions to the lightest shard: lists of session ids per shard,
    each in ascending session order."""
    load = [0] * shards
    out = [[] for _ in range(shards)]
    for k in sorted(ids, key=lambda k: -sizes[k]):
        j = int(np.argmin(load))
        out[j].append(k)
        load[j] += sizes[k]
    return [sorted(o) for o in out]


def poisson_arrivals(rate: float, duration: float, sessions, seed: int = 4):
    """(time, session) step requests: Poisson process of the given rate, sessions uniform."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = int(rng.poisson(rate * duration))
    t = np.sort(rng.uniform(0.0, duration, n))
    s = rng.choice(np.asarray(sessions), n)
    return list(zip(t.tolist(), s.tolist()))


@dataclass
class Request:
    session: int
    t_arrival: float
    t_done: float = -1.0


class Server:
    def __init__(self, scenes, params, budget_bytes: int = 200 << 30, shards: int = 8, graphs: bool = True,
                 max_tick: int = 0, device: int = 0):
        self.scenes = scenes
        self.params = params
        self.admitted, self.deferred, self.bytes_used = admit(scenes, budget_bytes)
        sizes = [sc.n for sc in scenes]
        groups = [g for g in shard_sessions(sizes, self.admitted, max(1, shards)) if g]
        self.shards = []
        self.where = {}
        for j, g in enumerate(groups):
            s = Solver(device)
            s.reserve(sum(sizes[k] for k in g), max(params.max_fluid_neighbors, 64))
            s.upload_batch([scenes[k] for k in g])
            s.graphs(graphs)
            self.shards.append(s)
            for local, k in enumerate(g):
                self.where[k] = (j, local)
        self.groups = groups
        self.max_tick = max_tick
        self.ticks = 0
        self.tick_sizes = []

    def close(self):
        for s in self.shards:
            if hasattr(s, "_tensor_views"):
                del s._tensor_views
            s.close()
        self.shards = []

    def session_state(self, k: int):
        """(pos, prev, vel) of session k, particle-index order (blocks until its shard is idle)."""
        j, local = self.where[k]
        self.shards[j].synchronize()
        return self.shards[j].readback_scene(local)

    def session_tensors(self, k: int):
        """Zero-copy torch views (positions_inv_mass, velocities; float32 (count, 4), particle-index
        order) of session k's resident state, and the shard's torch stream that writes them: read
        them after ``torch.cuda.current_stream().wait_stream(stream)`` once the tick that served the
        session is done.  The views alias the shard's buffers, so they follow later ticks."""
        j, local = self.where[k]
        s = self.shards[j]
        if not hasattr(s, "_tensor_views"):
            s._tensor_views = s.torch_views()
        first, count = s.ranges[local]
        v = s._tensor_views
        return (v["positions_inv_mass"][first:first + count], v["velocities"][first:first + count]), s.torch_stream()

    def reset_session(self, k: int, pos, vel):
        """Overwrite session k's particle state (an environment reset) without re-uploading its
        shard: ordered after the shard's in-flight tick on the shard's stream."""
        j, local = self.where[k]
        self.shards[j].set_scene_state(local, pos, vel)

    def serve(self, arrivals, warm: bool = True):
        """Serve (time, session) requests in real time; returns the completed Requests (deferred
        sessions' requests are refused and not returned)."""
        if warm:  # capture every shard's graph before the clock starts
            for s in self.shards:
                s.step_scenes(self.params, [])
            for s in self.shards:
                s.synchronize()
        pending = [collections.deque() for _ in self.shards]
        inflight = [None] * len(self.shards)
        done = []
        i = 0
        n = len(arrivals)
        t0 = time.perf_counter()
        remaining = sum(1 for _, k in arrivals if k in self.where)
        while remaining > 0:
            now = time.perf_counter() - t0
            while i < n and arrivals[i][0] <= now:
                t, k = arrivals[i]
                i += 1
                if k in self.where:
                    pending[self.where[k][0]].append(Request(k, t))
            busy = False
            for j, s in enumerate(self.shards):
                if inflight[j] is not None:
                    if not s.done():
                        busy = True
                        continue
                    tnow = time.perf_counter() - t0
                    for r in inflight[j]:
                        r.t_done = tnow
                    done.extend(inflight[j])
                    remaining -= len(inflight[j])
                    inflight[j] = None
                if pending[j]:
                    batch, seen, keep = [], set(), collections.deque()
                    while pending[j]:
                        r = pending[j].popleft()
                        if r.session in seen or (self.max_tick and len(batch) >= self.max_tick):
                            keep.append(r)
                        else:
                            seen.add(r.session)
                            batch.append(r)
                    pending[j] = keep
                    s.step_scenes(self.params, [self.where[r.session][1] for r in batch])
                    inflight[j] = batch
                    self.ticks += 1
                    self.tick_sizes.append(len(batch))
                    busy = True
            if not busy and i < n:
                wait = arrivals[i][0] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.001))
        self.elapsed = time.perf_counter() - t0
        return done


def latency_stats(done, elapsed: float):
    lat = np.array([r.t_done - r.t_arrival for r in done]) * 1e3
    return {"requests": len(done), "env_steps_per_s": len(done) / elapsed if elapsed > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
            "mean_ms": float(lat.mean()) if len(lat) else None}
