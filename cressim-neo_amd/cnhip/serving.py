"""Continuous-batching step server for many independent scenes resident in HBM.

Sessions are admitted against a device-memory budget and placed, balanced by particle count, in
a few shard pools.  Each shard is one context (one HIP stream) holding its sessions as a scene
pool.  The loop takes step requests as they arrive; whenever a shard is idle it steps, in one
launch per pass (``cn_step_scenes``, graph-captured), every session of the shard that has a pending
request, at most one request per session per tick so a session's steps stay in arrival order.
Requests that arrive while a tick runs wait for the next one; ticks of different shards run on
different streams and overlap on the device.  A request's latency runs from its arrival to the
host observing the completion of the tick that served it.

Membership changes as sessions come and go: a closed session is never stepped again, deferred
sessions are admitted into a new shard when the budget allows (a shard whose sessions all closed
returns its bytes; ``compact()`` rebuilds a sparse shard from its open sessions' state), and no
resident shard is uploaded again when another shard's membership changes.

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


class Shard:
    """One scene pool on its own context and stream: the sessions it holds (global ids, in pool
    order), which of them are still open, and the bytes its admission charged."""

    def __init__(self, solver: Solver, sessions, nbytes: int):
        self.solver = solver
        self.sessions = list(sessions)
        self.open = set(self.sessions)
        self.bytes = nbytes


class Server:
    """Continuous-batching server.  Sessions are admitted in order against the memory budget; those
    that do not fit wait (deferred) and are admitted as memory frees: a closed session stops being
    stepped at once, a shard whose sessions are all closed is destroyed and its bytes returned, and
    ``compact()`` rebuilds a sparse shard from its open sessions' current state.  Newly admitted
    sessions go into a new shard, so no resident shard is uploaded again when membership changes."""

    def __init__(self, scenes, params, budget_bytes: int = 200 << 30, shards: int = 8, graphs: bool = True,
                 max_tick: int = 0, device: int = 0):
        self.scenes = scenes
        self.params = params
        self.budget = budget_bytes
        self.graphs = graphs
        self.device = device
        self.max_tick = max_tick
        admitted, deferred, _ = admit(scenes, budget_bytes)
        self.deferred = collections.deque(deferred)
        self.closed = set()
        self.bytes_used = 0
        self.shards = []          # Shard or None (destroyed); indices stay stable
        self.where = {}           # open admitted session -> (shard index, index in the shard's pool)
        self.waiting = collections.defaultdict(collections.deque)  # requests of deferred sessions
        self.refused = 0
        self.admissions = []      # (session, shard index) in admission order
        sizes = [sc.n for sc in scenes]
        for g in shard_sessions(sizes, admitted, max(1, shards)):
            if g:
                self._add_shard(g)
        self.ticks = 0
        self.tick_sizes = []
        self._t0 = None           # perf_counter() at the start of the last serve() call

    # -- membership
    @property
    def admitted(self):
        return sorted(self.where)

    def _add_shard(self, sessions, states=None):
        sessions = list(sessions)
        nbytes = sum(estimate_bytes(self.scenes[k]) for k in sessions)
        sc = [self.scenes[k] if states is None else states[k] for k in sessions]
        s = Solver(self.device)
        s.reserve(sum(x.n for x in sc), max(self.params.max_fluid_neighbors, 64))
        s.upload_batch(sc)
        s.graphs(self.graphs)
        j = len(self.shards)
        self.shards.append(Shard(s, sessions, nbytes))
        self.bytes_used += nbytes
        for local, k in enumerate(sessions):
            self.where[k] = (j, local)
            self.admissions.append((k, j))
        return j

    def _drop_shard(self, j: int):
        sh = self.shards[j]
        if hasattr(sh.solver, "_tensor_views"):
            del sh.solver._tensor_views
        sh.solver.close()
        self.bytes_used -= sh.bytes
        self.shards[j] = None

    def close_session(self, k: int):
        """Ends session k: its queued requests are dropped, it is never stepped again, and when its
        shard has no open session left the shard goes and the deferred sessions that now fit are
        admitted."""
        if k in self.closed:
            return
        self.closed.add(k)
        if k in self.waiting:
            self.refused += len(self.waiting.pop(k))
        if k in self.deferred:
            self.deferred.remove(k)
            return
        j, _ = self.where.pop(k)
        sh = self.shards[j]
        sh.open.discard(k)
        if not sh.open:
            sh.solver.synchronize()
            self._drop_shard(j)
        self.admit_pending()

    def admit_pending(self) -> list:
        """Admits deferred sessions, in order, while they fit the budget, into one new shard;
        returns them."""
        batch, extra = [], 0
        while self.deferred:
            b = estimate_bytes(self.scenes[self.deferred[0]])
            if self.bytes_used + extra + b > self.budget:
                break
            batch.append(self.deferred.popleft())
            extra += b
        if batch:
            self._add_shard(batch)
        return batch

    def compact(self, min_closed: float = 0.5) -> int:
        """Rebuilds every shard whose closed share of sessions is at least min_closed from its open
        sessions' current state (read back from the device, uploaded as a new shard), returning the
        closed sessions' bytes to the budget, then admits what fits.  Returns the shards rebuilt."""
        rebuilt = 0
        for j, sh in enumerate(self.shards):
            if sh is None or not sh.open:
                continue
            if 1.0 - len(sh.open) / len(sh.sessions) < min_closed:
                continue
            keep = [k for k in sh.sessions if k in sh.open]
            states = {}
            for k in keep:
                pos, _, vel = self.session_state(k)
                sc = self.scenes[k].copy()
                sc.pos, sc.vel = pos.copy(), vel.copy()
                states[k] = sc
            for k in keep:
                del self.where[k]
            self._drop_shard(j)
            self._add_shard(keep, states)
            rebuilt += 1
        self.admit_pending()
        return rebuilt

    def close(self):
        for j, sh in enumerate(self.shards):
            if sh is not None:
                self._drop_shard(j)
        self.shards = []

    # -- state
    def session_state(self, k: int):
        """(pos, prev, vel) of session k, particle-index order (blocks until its shard is idle)."""
        j, local = self.where[k]
        s = self.shards[j].solver
        s.synchronize()
        return s.readback_scene(local)

    def session_tensors(self, k: int):
        """Zero-copy torch views (positions_inv_mass, velocities; float32 (count, 4), particle-index
        order) of session k's resident state, and the shard's torch stream that writes them: read
        them after ``torch.cuda.current_stream().wait_stream(stream)`` once the tick that served the
        session is done.  The views alias the shard's buffers, so they follow later ticks."""
        j, local = self.where[k]
        s = self.shards[j].solver
        if not hasattr(s, "_tensor_views"):
            s._tensor_views = s.torch_views()
        first, count = s.ranges[local]
        v = s._tensor_views
        return (v["positions_inv_mass"][first:first + count], v["velocities"][first:first + count]), s.torch_stream()

    def reset_session(self, k: int, pos, vel):
        """Overwrite session k's particle state (an environment reset) without re-uploading its
        shard: ordered after the shard's in-flight tick on the shard's stream."""
        j, local = self.where[k]
        self.shards[j].solver.set_scene_state(local, pos, vel)

    # -- serving
    def serve(self, arrivals, warm: bool = True, events=()):
        """Serve (time, session) requests in real time and return the completed Requests.  A
        deferred session's requests wait until it is admitted; a closed session's are refused (not
        returned).  events: (time, callable(server)) run once the clock passes their time (session
        closes, for instance); the call returns when every arrival and event has been handled and
        no admitted session has a request left (requests of sessions still deferred then stay in
        ``waiting``, and a later call serves them once their session is admitted: their arrival
        times move to that call's clock, before its start)."""
        if warm:  # capture every shard's graph before the clock starts
            for sh in self.shards:
                if sh is not None:
                    sh.solver.step_scenes(self.params, [])
            for sh in self.shards:
                if sh is not None:
                    sh.solver.synchronize()
        pending = collections.defaultdict(collections.deque)
        inflight = {}
        done = []
        events = sorted(events, key=lambda e: e[0])
        i = ie = 0
        n = len(arrivals)
        t0 = time.perf_counter()
        if self._t0 is not None:
            shift = t0 - self._t0
            for q in self.waiting.values():
                for r in q:
                    r.t_arrival -= shift
        self._t0 = t0
        while True:
            now = time.perf_counter() - t0
            while ie < len(events) and events[ie][0] <= now:
                events[ie][1](self)
                ie += 1
            # requests of sessions admitted since they arrived
            for k in [k for k in self.waiting if k in self.where]:
                pending[self.where[k][0]].extend(self.waiting.pop(k))
            while i < n and arrivals[i][0] <= now:
                t, k = arrivals[i]
                i += 1
                if k in self.closed:
                    self.refused += 1
                elif k in self.where:
                    pending[self.where[k][0]].append(Request(k, t))
                else:
                    self.waiting[k].append(Request(k, t))
            busy = False
            for j, sh in enumerate(self.shards):
                if sh is None:  # dropped: its last tick finished before it went
                    if j in inflight:
                        tnow = time.perf_counter() - t0
                        for r in inflight[j]:
                            r.t_done = tnow
                        done.extend(inflight.pop(j))
                    if pending.get(j):
                        self.refused += len(pending.pop(j))
                    continue
                s = sh.solver
                if j in inflight:
                    if not s.done():
                        busy = True
                        continue
                    tnow = time.perf_counter() - t0
                    for r in inflight[j]:
                        r.t_done = tnow
                    done.extend(inflight.pop(j))
                q = pending.get(j)
                if q:
                    batch, seen, keep = [], set(), collections.deque()
                    while q:
                        r = q.popleft()
                        if r.session not in sh.open:
                            self.refused += 1  # closed while it waited
                        elif r.session in seen or (self.max_tick and len(batch) >= self.max_tick):
                            keep.append(r)
                        else:
                            seen.add(r.session)
                            batch.append(r)
                    pending[j] = keep
                    if batch:
                        s.step_scenes(self.params, [self.where[r.session][1] for r in batch])
                        inflight[j] = batch
                        self.ticks += 1
                        self.tick_sizes.append(len(batch))
                        busy = True
            work_left = inflight or any(pending.values())
            if i >= n and ie >= len(events) and not work_left:
                break
            if not busy:
                nxt = [arrivals[i][0]] if i < n else []
                if ie < len(events):
                    nxt.append(events[ie][0])
                if nxt:
                    wait = min(nxt) - (time.perf_counter() - t0)
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
