"""Scenes of more than 32 static colliders (the grouped candidate search and the per-slot collider
lists): a fluid block and a cloth on a board of 60 pegs steps bit for bit like the oracle, alone, in
a graph-captured pool beside scenes of few colliders, through compact ticks and after a wall is
moved; pegs out of reach change nothing (the list path equals the mask path of the same scene
without them)."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

pytestmark = pytest.mark.gpu


def params(**kw):
    base = dict(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    base.update(kw)
    return cnhip.default_params(**base)


def with_far_pegs(sc):
    alt = sc.copy()
    far = scenes.pegboard_colliders(8, 5, y=1.6, origin=(2.2, 0.3))
    alt.shape_colliders = np.concatenate([far[:20], sc.shape_colliders, far[20:]])
    alt.name = sc.name + "_far_pegs"
    return alt


def test_pegboard_bitexact(solver, oracle):
    sc = scenes.pegboard()
    assert len(sc.colliders) + len(sc.shape_colliders) > 32
    p = params()
    osc = oracle.OracleScene(sc)
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    for step in range(28):   # the fluid falls 0.1 m onto the pegs in about 17 steps
        print(f"pegboard step {step}", flush=True)
        solver.step(p)
        oracle.step(osc, p)
        pos, _, vel = solver.readback()
        np.testing.assert_array_equal(pos, osc.pos, err_msg=f"step {step}")
        np.testing.assert_array_equal(vel, osc.vel, err_msg=f"step {step}")
    assert solver.meta().contact_overflows == 0
    closest = np.full(sc.n, np.inf)
    for col in sc.shape_colliders:
        closest = np.minimum(closest, [oracle.collider_contact(col, q)[1] for q in osc.pos[:, :3]])
    assert (closest < 1.1 * float(sc.radii[0])).sum() > 20   # the pegs were reached


def test_out_of_reach_pegs_change_nothing(solver, oracle):
    sc = scenes.shapes_scene(side=8)
    alt = with_far_pegs(sc)
    p = params(substeps=2, default_iterations=3)
    got = []
    for x in (sc, alt):
        print(f"{x.name}: {len(x.colliders) + len(x.shape_colliders)} colliders", flush=True)
        solver.reserve(x.n, 64)
        solver.upload(x)
        for _ in range(15):
            solver.step(p)
        got.append(solver.readback())
    osc = oracle.OracleScene(alt)
    for _ in range(15):
        oracle.step(osc, p)
    for a, b in zip(got[0], got[1]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(got[1][0], osc.pos)
    np.testing.assert_array_equal(got[1][2], osc.vel)


def pool_batch():
    return [scenes.batch_scene(1800, 0, 1), scenes.pegboard(fluid_side=8, cloth_n=10), scenes.shapes_scene(side=8),
            scenes.batch_scene(1500, 2, 3), with_far_pegs(scenes.shapes_scene(side=6)), scenes.batch_scene(1200, 1, 4)]


def test_pegboard_in_a_graph_pool(oracle):
    batch = pool_batch()
    p = params(substeps=2)
    steps = 20
    with cnhip.Solver(0) as pool:
        pool.reserve(sum(sc.n for sc in batch), 64)
        pool.upload_batch(batch)
        pool.graphs(True)
        for s in range(steps):
            pool.step(p)
        got = [pool.readback_scene(k) for k in range(len(batch))]
        assert pool.meta().contact_overflows == 0
    for k, sc in enumerate(batch):
        print(f"oracle scene {k} {sc.name}", flush=True)
        osc = oracle.OracleScene(sc)
        for _ in range(steps):
            oracle.step(osc, p)
        np.testing.assert_array_equal(got[k][0], osc.pos, err_msg=f"scene {k} {sc.name}")
        np.testing.assert_array_equal(got[k][2], osc.vel, err_msg=f"scene {k} {sc.name}")


def test_pegboard_through_compact_ticks_and_a_moved_wall(solver, oracle):
    """Compact ticks over the pegboard and its neighbours equal the scenes stepped alone; then the
    pegboard's x-low wall moves inward (cn_scene_set_colliders, the groups rebuilt) and the pool
    still steps like the oracle."""
    batch = pool_batch()
    p = params(substeps=2)
    seq = [None, [1], [1, 4], None, [0, 1], [4]]
    counts = np.zeros(len(batch), int)
    with cnhip.Solver(0) as pool:
        pool.reserve(sum(sc.n for sc in batch), 64)
        pool.upload_batch(batch)
        pool.tick_policy(0.5)
        for t, ids in enumerate(seq):
            print(f"tick {t} scenes {ids}", flush=True)
            if ids is None:
                pool.step(p)
                counts += 1
            else:
                pool.step_scenes(p, ids)
                counts[ids] += 1
        got = [pool.readback_scene(k) for k in range(len(batch))]
        oscs = []
        for k, sc in enumerate(batch):
            osc = oracle.OracleScene(sc)
            for _ in range(counts[k]):
                oracle.step(osc, p)
            np.testing.assert_array_equal(got[k][0], osc.pos, err_msg=f"scene {k} {sc.name} ({counts[k]} steps)")
            np.testing.assert_array_equal(got[k][2], osc.vel, err_msg=f"scene {k} {sc.name}")
            oscs.append(osc)
        new = batch[1].colliders.copy()
        new["center"][0] += np.float32(0.19)   # the wall face meets the fluid block
        pool.set_scene_colliders(1, new)
        oscs[1].colliders[:] = new
        for s in range(3):
            print(f"after the wall moved: step {s}", flush=True)
            pool.step(p)
            for o in oscs:
                oracle.step(o, p)
        for k, o in enumerate(oscs):
            pos, _, vel = pool.readback_scene(k)
            np.testing.assert_array_equal(pos, o.pos, err_msg=f"scene {k} after the edit")
            np.testing.assert_array_equal(vel, o.vel, err_msg=f"scene {k} after the edit")
