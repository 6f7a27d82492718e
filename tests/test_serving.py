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
