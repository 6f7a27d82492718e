"""GPU parity of static colliders of any shape: a fluid block falling on a sphere, a capsule and an
oriented box steps bit for bit like the CPU oracle, alone and in a pool, and the plain box walls
given as shape colliders step like the plain boxes."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

pytestmark = pytest.mark.gpu


def params(**kw):
    base = dict(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    base.update(kw)
    return cnhip.default_params(**base)


def test_shapes_scene_bitexact(solver, oracle):
    sc = scenes.shapes_scene(side=12)
    p = params()
    osc = oracle.OracleScene(sc)
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    for step in range(8):
        print(f"shapes_scene(12) step {step}", flush=True)
        solver.step(p)
        oracle.step(osc, p)
        pos, _, vel = solver.readback()
        np.testing.assert_array_equal(pos, osc.pos, err_msg=f"step {step}")
        np.testing.assert_array_equal(vel, osc.vel, err_msg=f"step {step}")
    assert solver.meta().fluid_neighbor_truncations == 0


def test_shape_scenes_in_a_pool(oracle):
    batch = [scenes.shapes_scene(side=8, seed=1), scenes.batch_scene(1500, 0, seed=2), scenes.shapes_scene(side=10)]
    p = params(substeps=2)
    with cnhip.Solver(0) as pool:
        pool.reserve(sum(sc.n for sc in batch), 64)
        pool.upload_batch(batch)
        for _ in range(3):
            pool.step(p)
        got = [pool.readback_scene(k) for k in range(len(batch))]
    for k, sc in enumerate(batch):
        osc = oracle.OracleScene(sc)
        for _ in range(3):
            oracle.step(osc, p)
        np.testing.assert_array_equal(got[k][0], osc.pos, err_msg=f"scene {k}")
        np.testing.assert_array_equal(got[k][2], osc.vel, err_msg=f"scene {k}")


def test_walls_as_shape_colliders_equal_plain_boxes(solver):
    sc = scenes.dam_break_64k(side=12)
    alt = sc.copy()
    alt.shape_colliders = np.concatenate([
        cnhip.shape_collider(1, tuple(b["center"]), params=tuple(b["half_extents"]), material=tuple(b["material"]))
        for b in sc.colliders])
    alt.colliders = sc.colliders[:0]
    p = params()
    out = []
    for s in (sc, alt):
        solver.reserve(s.n, 64)
        solver.upload(s)
        for _ in range(4):
            solver.step(p)
        out.append(solver.readback())
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
