"""Compact ticks (cn_tick_policy): cn_step_scenes over a small part of a pool runs on the listed
scenes' slots alone.  A pool of fluid, soft-body, cloth, contact and shape scenes is stepped
through a seeded sequence of ticks of varying membership, whole-pool steps in between, once with
compact ticks and once with every tick over the whole pool: both pools end bit for bit equal, and
equal to each scene stepped alone as often as it was listed."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

pytestmark = pytest.mark.gpu


def params():
    return cnhip.default_params(dt=1 / 120, substeps=2, default_iterations=3, max_fluid_neighbors=64)


def pool_scenes():
    sh = scenes.shapes_scene(side=8)
    return [scenes.batch_scene(2200, 0, 1), scenes.batch_scene(1500, 1, 2), scenes.batch_scene(1800, 2, 3),
            scenes.contact_stack(block_res=4, cloth_n=12, fluid_side=4), sh, scenes.batch_scene(3000, 0, 4),
            scenes.batch_scene(1200, 1, 5), scenes.folded_cloth(n=12)]


def ticks(n_scenes, seed=5):
    """A seeded sequence of ticks: small subsets (compact), an empty tick, whole-pool steps (None)."""
    rng = np.random.default_rng(seed)
    seq = []
    for k in range(10):
        if k in (3, 7):
            seq.append(None)
        elif k == 5:
            seq.append([])
        else:
            size = int(rng.integers(1, 4))
            seq.append(sorted(rng.choice(n_scenes, size, replace=False).tolist()))
    return seq


def run_pool(batch, seq, compact_below):
    p = params()
    counts = np.zeros(len(batch), int)
    with cnhip.Solver(0) as pool:
        pool.reserve(sum(sc.n for sc in batch), 64)
        pool.upload_batch(batch)
        pool.tick_policy(compact_below)
        pool.step(p)  # one whole-pool step first: the slot order then holds every scene as one segment
        counts += 1
        for t, ids in enumerate(seq):
            print(f"compact_below {compact_below}: tick {t} scenes {ids}", flush=True)
            if ids is None:
                pool.step(p)
                counts += 1
            else:
                pool.step_scenes(p, ids)
                counts[ids] += 1
        pool.synchronize()
        out = [pool.readback_scene(k) for k in range(len(batch))]
        meta = pool.meta()
    return out, counts, meta


def test_compact_ticks_equal_whole_pool_ticks_and_scenes_alone(solver):
    batch = pool_scenes()
    seq = ticks(len(batch))
    a, counts, meta = run_pool(batch, seq, 0.5)
    b, counts_b, _ = run_pool(batch, seq, 0.0)
    assert (counts == counts_b).all()
    assert meta.fluid_neighbor_truncations == 0
    for k in range(len(batch)):
        for arr_a, arr_b in zip(a[k], b[k]):
            np.testing.assert_array_equal(arr_a, arr_b, err_msg=f"scene {k} {batch[k].name}")
    p = params()
    for k in range(len(batch)):
        solver.reserve(batch[k].n, 64)
        solver.upload(batch[k])
        for _ in range(counts[k]):
            solver.step(p)
        pos, prev, vel = solver.readback()
        np.testing.assert_array_equal(a[k][0], pos, err_msg=f"scene {k} alone ({counts[k]} steps)")
        np.testing.assert_array_equal(a[k][2], vel, err_msg=f"scene {k} alone")


def test_compact_ticks_before_any_whole_pool_step_and_with_graphs():
    """Before a whole-pool step has sorted the pool, a tick steps the whole pool (the slots still
    hold the upload order); with graph capture on, compact ticks and captured whole-pool steps mix."""
    batch = pool_scenes()[:5]
    p = params()
    results = []
    for compact in (0.6, 0.0):
        with cnhip.Solver(0) as pool:
            pool.reserve(sum(sc.n for sc in batch), 64)
            pool.upload_batch(batch)
            pool.graphs(True)
            pool.tick_policy(compact)
            pool.step_scenes(p, [1])
            pool.step(p)
            pool.step_scenes(p, [0, 3])
            pool.step_scenes(p, [4])
            pool.step(p)
            pool.step_scenes(p, [2, 4])
            results.append([pool.readback_scene(k) for k in range(len(batch))])
    for k in range(len(batch)):
        for x, y in zip(results[0][k], results[1][k]):
            np.testing.assert_array_equal(x, y, err_msg=f"scene {k}")
