"""The scene pool: scenes uploaded together and stepped with one launch per pass give, scene by
scene, the bits of stepping each scene alone (and of the CPU oracle)."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

pytestmark = pytest.mark.gpu


def params(**kw):
    base = dict(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    base.update(kw)
    return cnhip.default_params(**base)


def small_batch():
    sizes = [(3000, 0), (900, 1), (1600, 2), (5000, 0), (2200, 1), (400, 2), (1200, 0)]
    return [scenes.batch_scene(n, k, seed=i) for i, (n, k) in enumerate(sizes)]


def test_batch_equals_scenes_alone(solver):
    batch = small_batch()
    p = params()
    with cnhip.Solver(0) as pool:
        pool.reserve(sum(sc.n for sc in batch), 64)
        pool.upload_batch(batch)
        assert [c for _, c in pool.ranges] == [sc.n for sc in batch]
        for _ in range(3):
            pool.step(p)
        got = [pool.readback_scene(k) for k in range(len(batch))]
        assert pool.meta().fluid_neighbor_truncations == 0
    for k, sc in enumerate(batch):
        solver.reserve(sc.n, 64)
        solver.upload(sc)
        for _ in range(3):
            solver.step(p)
        pos, prev, vel = solver.readback()
        np.testing.assert_array_equal(got[k][0], pos, err_msg=f"scene {k} positions")
        np.testing.assert_array_equal(got[k][2], vel, err_msg=f"scene {k} velocities")


def test_batch_matches_oracle(oracle):
    batch = small_batch()[:4]
    p = params()
    with cnhip.Solver(0) as pool:
        pool.reserve(sum(sc.n for sc in batch), 64)
        pool.upload_batch(batch)
        for _ in range(2):
            pool.step(p)
        got = [pool.readback_scene(k) for k in range(len(batch))]
    for k, sc in enumerate(batch):
        osc = oracle.OracleScene(sc)
        for _ in range(2):
            oracle.step(osc, p)
        np.testing.assert_array_equal(got[k][0], osc.pos, err_msg=f"scene {k}")
        np.testing.assert_array_equal(got[k][2], osc.vel, err_msg=f"scene {k}")


def test_batch_rejects_hash_keys():
    with cnhip.Solver(0) as pool:
        pool.upload_batch(small_batch()[:2])
        with pytest.raises(cnhip.CnError):
            pool.step(params(key_mode=cnhip.CN_KEY_HASH))


@pytest.mark.parametrize("substeps", [4, 3])
def test_graph_replays_equal_eager_steps(substeps):
    """cn_step captured into a HIP graph and replayed gives the eager step's bits on every replay,
    with eager work and synchronizations between the replays (odd substep counts alternate the
    slot-order parity, so two graphs are cached and used in turn)."""
    batch = small_batch()[:5]
    p = params(substeps=substeps)
    with cnhip.Solver(0) as eager, cnhip.Solver(0) as graphed:
        for s in (eager, graphed):
            s.reserve(sum(sc.n for sc in batch), 64)
            s.upload_batch(batch)
        graphed.graphs(True)
        for step in range(5):
            eager.step(p)
            graphed.step(p)
            a = eager.readback()
            b = graphed.readback()
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x, y, err_msg=f"step {step}")
            eager.meta()
            graphed.meta()


@pytest.mark.parametrize("graphs", [False, True])
def test_step_scenes_leaves_others_untouched(solver, graphs):
    """cn_step_scenes advances only the listed scenes; each scene ends where stepping it alone the
    same number of times ends, bit for bit."""
    batch = small_batch()
    p = params()
    ticks = [[0, 2, 4], [1], [0, 1, 2, 3, 4, 5, 6], [], [6, 3], [0]]
    counts = np.zeros(len(batch), int)
    with cnhip.Solver(0) as pool:
        pool.reserve(sum(sc.n for sc in batch), 64)
        pool.upload_batch(batch)
        pool.graphs(graphs)
        for t in ticks:
            pool.step_scenes(p, t)
            counts[t] += 1
        got = [pool.readback_scene(k) for k in range(len(batch))]
    for k, sc in enumerate(batch):
        solver.reserve(sc.n, 64)
        solver.upload(sc)
        for _ in range(counts[k]):
            solver.step(p)
        pos, prev, vel = solver.readback()
        np.testing.assert_array_equal(got[k][0], pos, err_msg=f"scene {k}")
        np.testing.assert_array_equal(got[k][1], prev, err_msg=f"scene {k} prev")
        np.testing.assert_array_equal(got[k][2], vel, err_msg=f"scene {k}")
