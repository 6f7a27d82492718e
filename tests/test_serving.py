"""The serving loop: admission by memory budget, shard balance, and (GPU) that every session's
state after being served equals stepping that scene alone as many times as it was served."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes, serving


def small_sessions():
    sizes = [(1500, 0), (700, 1), (900, 2), (2500, 0), (600, 1), (400, 2), (1100, 0), (800, 2)]
    return [scenes.batch_scene(n, k, seed=i) for i, (n, k) in enumerate(sizes)]


def test_admission_respects_budget():
    ses = small_sessions()
    b = [serving.estimate_bytes(sc) for sc in ses]
    budget = b[0] + b[1] + b[3] - 1  # sessions 0 and 1 fit; 2 fits after them only if small enough
    adm, dfr, used = serving.admit(ses, budget)
    assert used <= budget
    assert sorted(adm + dfr) == list(range(len(ses)))
    assert adm[:2] == [0, 1]
    assert sum(b[k] for k in adm) == used


def test_shards_balance_and_cover():
    sizes = [5000, 4000, 3000, 2000, 1000, 500, 250, 100]
    groups = serving.shard_sessions(sizes, list(range(8)), 3)
    assert sorted(k for g in groups for k in g) == list(range(8))
    loads = [sum(sizes[k] for k in g) for g in groups]
    assert max(loads) - min(loads) <= max(sizes)


def test_poisson_arrivals_rate_and_order():
    arr = serving.poisson_arrivals(1000.0, 2.0, range(16), seed=4)
    assert abs(len(arr) - 2000) < 5 * np.sqrt(2000)
    t = [a[0] for a in arr]
    assert t == sorted(t) and 0.0 <= t[0] and t[-1] <= 2.0


@pytest.mark.gpu
def test_served_sessions_equal_scenes_alone(solver):
    ses = small_sessions()
    p = cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    srv = serving.Server(ses, p, budget_bytes=1 << 40, shards=3, graphs=True)
    try:
        arr = serving.poisson_arrivals(400.0, 0.1, range(len(ses)), seed=7)
        done = srv.serve(arr)
        assert len(done) == len(arr)
        served = np.bincount([r.session for r in done], minlength=len(ses))
        states = [srv.session_state(k) for k in range(len(ses))]
    finally:
        srv.close()
    for k, sc in enumerate(ses):
        solver.reserve(sc.n, 64)
        solver.upload(sc)
        for _ in range(served[k]):
            solver.step(p)
        pos, _, vel = solver.readback()
        np.testing.assert_array_equal(states[k][0], pos, err_msg=f"session {k} ({served[k]} steps)")
        np.testing.assert_array_equal(states[k][2], vel, err_msg=f"session {k}")


def test_admission_queue_and_close_without_gpu():
    """The admission bookkeeping of a server: greedy admission in session order, the deferred
    queue, closes of deferred sessions (CPU only: no shard is built)."""
    ses = small_sessions()
    b = [serving.estimate_bytes(sc) for sc in ses]
    adm, dfr, used = serving.admit(ses, sum(b[:6]))
    assert adm == [0, 1, 2, 3, 4, 5] and dfr == [6, 7] and used == sum(b[:6])


@pytest.mark.gpu
def test_deferred_sessions_admitted_as_memory_frees(solver):
    """Sessions beyond the budget wait; closing every session of one shard drops that shard,
    returns its bytes and admits the waiting sessions into a new shard (the other shards stay as
    they are); the waiting sessions' requests are then served, and every open session ends where
    stepping it alone as often as it was served ends."""
    ses = small_sessions()
    p = cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    b = [serving.estimate_bytes(sc) for sc in ses]
    srv = serving.Server(ses, p, budget_bytes=sum(b[:6]), shards=3, graphs=True)
    try:
        assert list(srv.deferred) == [6, 7] and srv.admitted == [0, 1, 2, 3, 4, 5]
        first = srv.shards[0]
        gone = list(first.sessions)
        others = [id(sh.solver) for sh in srv.shards[1:]]
        arr = serving.poisson_arrivals(600.0, 0.15, range(len(ses)), seed=9)
        done = srv.serve(arr, events=[(0.05, lambda s: [s.close_session(k) for k in gone])])
        assert srv.shards[0] is None and 6 in srv.where
        assert [id(sh.solver) for sh in srv.shards[1:3]] == others  # untouched
        assert srv.bytes_used <= srv.budget
        served = np.bincount([r.session for r in done], minlength=len(ses))
        assert served[6] > 0
        for r in done:
            assert r.t_done >= r.t_arrival
        open_ = [k for k in range(len(ses)) if k in srv.where]
        states = {k: srv.session_state(k) for k in open_}
        # compact: close one of two sessions in a shard, rebuild it from the survivor's state
        two = next(sh for sh in srv.shards if sh is not None and len(sh.open) >= 2)
        k_close, k_keep = sorted(two.open)[:2]
        srv.close_session(k_close)
        assert srv.compact(min_closed=0.3) >= 1
        more = serving.poisson_arrivals(300.0, 0.05, [k_keep], seed=11)
        done2 = srv.serve(more)
        # every new request, plus those a session admitted by the compaction kept waiting
        assert sum(r.session == k_keep for r in done2) == len(more)
        for r in done2:
            assert r.t_done >= r.t_arrival
            served[r.session] += 1
        open_ = [k for k in range(len(ses)) if k in srv.where]
        states = {k: srv.session_state(k) for k in open_}
    finally:
        srv.close()
    for k in open_:
        solver.reserve(ses[k].n, 64)
        solver.upload(ses[k])
        for _ in range(served[k]):
            solver.step(p)
        pos, _, vel = solver.readback()
        np.testing.assert_array_equal(states[k][0], pos, err_msg=f"session {k} ({served[k]} steps)")
        np.testing.assert_array_equal(states[k][2], vel, err_msg=f"session {k}")
