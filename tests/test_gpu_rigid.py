"""GPU parity of particle-rigid contacts (solid particles against static colliders): a tet block
and a cloth falling onto the floor and onto a sphere, a capsule and an oriented box, the
single-particle floor, friction and restitution cases, and a mixed scene with fluid, a soft body
and a cloth step bit for bit like the CPU oracle, alone and in a graph-captured pool."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

pytestmark = pytest.mark.gpu


def params(**kw):
    base = dict(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    base.update(kw)
    return cnhip.default_params(**base)


def cloth_over_obstacles(n=14):
    part = scenes.cloth_part(n, 0.04, 0.02, translation=(0.3, 0.35, 0.35))
    sc = scenes.compose([part], cnhip.fluid_material(0.03), scenes.container((0, 0, 0), (4, 2, 2)),
                        f"cloth_obstacles_{n}")
    sc.shape_colliders = scenes.obstacles()
    return sc


def block_on_obstacles(res=6):
    part = scenes.soft_block_part(res, 0.04, 0.02, translation=(0.3, 0.45, 0.35))
    sc = scenes.compose([part], cnhip.fluid_material(0.03), scenes.container((0, 0, 0), (4, 2, 2)),
                        f"block_obstacles_{res}")
    sc.shape_colliders = scenes.obstacles()
    return sc


def single(y0, vel, r=0.02, floor_mat=(0.0, 0.0, 0.0, 0.0), part_mat=(0.0, 0.0, 0.0, 0.0)):
    part = scenes._solid_part(np.array([(0.1, y0, 0.2)], np.float32), r, 1.0)
    cols = scenes.container((0, 0, 0), (4, 2, 2), material=floor_mat)
    sc = scenes.compose([part], cnhip.fluid_material(0.03), cols, "single_rigid")
    sc.vel[:, :3] = np.asarray(vel, np.float32)
    sc.part_mats = np.array([part_mat], np.float32)
    return sc


def step_both(solver, oracle, sc, p, steps, label):
    osc = oracle.OracleScene(sc)
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    for step in range(steps):
        print(f"{label} step {step}", flush=True)
        solver.step(p)
        oracle.step(osc, p)
        pos, prev, vel = solver.readback()
        np.testing.assert_array_equal(pos, osc.pos, err_msg=f"{label} step {step}")
        np.testing.assert_array_equal(vel, osc.vel, err_msg=f"{label} step {step}")
    return osc


@pytest.mark.parametrize("case", ["block_floor", "cloth_floor", "block_obstacles", "cloth_obstacles"])
def test_solids_against_static_colliders_bitexact(solver, oracle, case):
    sc = {"block_floor": lambda: scenes.batch_scene(2000, 1, 0),
          "cloth_floor": lambda: scenes.batch_scene(2000, 2, 0),
          "block_obstacles": block_on_obstacles,
          "cloth_obstacles": cloth_over_obstacles}[case]()
    # the block falls 0.05 m, the cloth 1.15 m: enough steps for the floor contacts to act
    steps = 30 if case == "cloth_floor" else 12
    osc = step_both(solver, oracle, sc, params(), steps, case)
    assert np.isfinite(osc.pos).all()
    assert osc.pos[:, 1].min() > 0.0


@pytest.mark.parametrize("case", ["floor", "restitution", "friction_static", "friction_kinetic"])
def test_single_particle_cases_bitexact(solver, oracle, case):
    p = params(substeps=1, default_iterations=2, gravity=(0.0, 0.0 if case != "floor" else -9.81, 0.0, 0.0))
    sc = {"floor": lambda: single(0.021, (0.0, -0.5, 0.0)),
          "restitution": lambda: single(0.02 - 0.0015 + 2.0 / 120, (0.0, -2.0, 0.0), floor_mat=(0.0, 0.6, 0.0, 0.0),
                                        part_mat=(0.0, 0.2, 0.0, 0.0)),
          "friction_static": lambda: single(0.016, (0.1, 0.0, 0.05), floor_mat=(0.5, 0.0, 0.0, 1.0),
                                            part_mat=(0.5, 0.0, 0.0, 1.0)),
          "friction_kinetic": lambda: single(0.016, (1.2, 0.0, -0.4), floor_mat=(0.5, 0.0, 0.0, 1.0),
                                             part_mat=(0.5, 0.0, 0.0, 1.0))}[case]()
    step_both(solver, oracle, sc, p, 3, case)


def test_mixed_scene_with_walls_bitexact(solver, oracle):
    """Fluid, a tet block and a cloth in one container: fluid boundary, particle-particle and
    particle-rigid contacts in the same contact iterations."""
    sc = scenes.contact_stack(block_res=5, cloth_n=16, fluid_side=5)
    sc.shape_colliders = scenes.obstacles()
    step_both(solver, oracle, sc, params(), 6, "contact_stack")


def test_rigid_scenes_in_a_graph_pool(oracle):
    batch = [scenes.batch_scene(1800, 1, 1), scenes.batch_scene(2200, 0, 2), cloth_over_obstacles(10),
             scenes.batch_scene(2500, 2, 3), block_on_obstacles(5)]
    p = params(substeps=2)
    with cnhip.Solver(0) as pool:
        pool.reserve(sum(sc.n for sc in batch), 64)
        pool.upload_batch(batch)
        pool.graphs(True)
        for _ in range(4):
            pool.step(p)
        got = [pool.readback_scene(k) for k in range(len(batch))]
    for k, sc in enumerate(batch):
        osc = oracle.OracleScene(sc)
        for _ in range(4):
            oracle.step(osc, p)
        np.testing.assert_array_equal(got[k][0], osc.pos, err_msg=f"scene {k} {sc.name}")
        np.testing.assert_array_equal(got[k][2], osc.vel, err_msg=f"scene {k} {sc.name}")
