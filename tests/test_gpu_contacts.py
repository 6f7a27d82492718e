"""GPU parity of particle-particle contacts (candidate lists, Jacobi position solve, velocity solve)
against the CPU oracle, alone and inside a scene pool, self collision of a folded cloth, plus the
capacity path."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

pytestmark = pytest.mark.gpu


def params(iters=4, substeps=4, **kw):
    return cnhip.default_params(dt=1 / 120, substeps=substeps, default_iterations=iters, max_fluid_neighbors=64,
                                **kw)


def run_pair(sc, p, steps, solver, oracle):
    osc = oracle.OracleScene(sc)
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    out = []
    for _ in range(steps):
        solver.step(p)
        st = oracle.step(osc, p)
        m = solver.meta()
        assert m.contact_overflows == st.contact_overflows
        assert m.max_cell_occupancy > 0
        pos, _, vel = solver.readback()
        assert np.isfinite(pos).all()
        out.append((float(np.abs(pos - osc.pos).max()), float(np.abs(vel - osc.vel).max()),
                    bool((pos == osc.pos).all() and (vel == osc.vel).all())))
    return out


@pytest.mark.parametrize("iters,substeps", [(1, 1), (3, 2), (5, 4)])
def test_contact_stack_bitexact(solver, oracle, iters, substeps):
    print(f"contact_stack(4, 12, 3) iters={iters} substeps={substeps}", flush=True)
    res = run_pair(scenes.contact_stack(4, 12, 3), params(iters=iters, substeps=substeps), 5, solver, oracle)
    print(res)
    assert all(b for _, _, b in res), res


def test_contact_stack_contact_iterations_differ(solver, oracle):
    print("contact_stack(6, 24, 6) contact iterations 2, fluid 3, soft 5", flush=True)
    p = params(iters=3, soft_contact_iterations=2, soft_internal_iterations=5)
    res = run_pair(scenes.contact_stack(6, 24, 6), p, 8, solver, oracle)
    print(res)
    assert all(b for _, _, b in res), res


def test_contact_stack_large(solver, oracle):
    print("contact_stack(12, 64, 16)", flush=True)
    res = run_pair(scenes.contact_stack(12, 64, 16), params(), 4, solver, oracle)
    print(res)
    assert all(b for _, _, b in res), res


def test_contact_overflow_is_reported(solver, oracle):
    rng = np.random.default_rng(0)
    a = scenes._solid_part(np.zeros((1, 3), np.float32), 0.01, 1.0)
    b = scenes._solid_part(rng.uniform(-0.004, 0.004, (40, 3)).astype(np.float32), 0.01, 1.0)
    sc = scenes.compose([a, b], cnhip.fluid_material(0.03), np.zeros(0, cnhip.COLLIDER_DTYPE), "overflow")
    sc.pos[:, 3] = 0.0
    p = params(iters=2, substeps=3)
    res = run_pair(sc, p, 2, solver, oracle)
    assert all(b for _, _, b in res), res
    assert solver.meta().contact_overflows == 3
    with pytest.raises(cnhip.CnError) as e:
        solver.meta(check=True)
    assert e.value.status == cnhip.abi.CN_ERR_CAPACITY


@pytest.mark.parametrize("iters,substeps", [(1, 1), (4, 4)])
def test_self_colliding_cloth_fold_bitexact(solver, oracle, iters, substeps):
    """A cloth folded onto itself with self collision: the layers push apart where they are not
    constraint-adjacent (IsAdjacentParticle keeps the hinge opposites across the crease out), bit
    for bit against the oracle; without the flag the layers pass through each other."""
    print(f"folded_cloth(16) iters={iters} substeps={substeps}", flush=True)
    p = params(iters=iters, substeps=substeps)
    res = run_pair(scenes.folded_cloth(16), p, 4, solver, oracle)
    print(res)
    assert all(b for _, _, b in res), res
    collided = solver.readback()[0]
    res = run_pair(scenes.folded_cloth(16, self_collide=False), p, 4, solver, oracle)
    assert all(b for _, _, b in res), res
    assert not np.array_equal(collided, solver.readback()[0])


def test_self_collision_in_a_pool_equals_alone(solver, oracle):
    batch = [scenes.folded_cloth(12), scenes.contact_stack(4, 12, 3), scenes.folded_cloth(20)]
    p = params(iters=3, substeps=2)
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


def test_contact_scenes_in_a_pool_equal_alone(solver, oracle):
    batch = [scenes.contact_stack(4, 12, 3), scenes.batch_scene(1500, 0, seed=1), scenes.contact_stack(6, 24, 6),
             scenes.batch_scene(900, 2, seed=2), scenes.contact_stack(5, 16, 4)]
    p = params(iters=3, substeps=2)
    with cnhip.Solver(0) as pool:
        pool.reserve(sum(sc.n for sc in batch), 64)
        pool.upload_batch(batch)
        pool.step(p)
        pool.step_scenes(p, [0, 2])  # the others keep their state
        pool.synchronize()
        got = [pool.readback_scene(k) for k in range(len(batch))]
    for k, sc in enumerate(batch):
        osc = oracle.OracleScene(sc)
        for _ in range(2 if k in (0, 2) else 1):
            oracle.step(osc, p)
        np.testing.assert_array_equal(got[k][0], osc.pos, err_msg=f"scene {k}")
        np.testing.assert_array_equal(got[k][2], osc.vel, err_msg=f"scene {k}")
