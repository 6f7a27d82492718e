"""Per-scene state overwrite (cn_scene_set_state): an environment reset in a pool without a
re-upload steps bit for bit like the scene uploaded alone, leaves the other scenes untouched, and
takes host arrays or torch tensors on the device."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

pytestmark = pytest.mark.gpu


def params():
    return cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)


def alone(sc, steps):
    s = cnhip.Solver(0)
    try:
        s.reserve(sc.n, 64)
        s.upload(sc)
        for _ in range(steps):
            s.step(params())
        return s.readback()
    finally:
        s.close()


def test_reset_scene_in_pool_steps_like_alone():
    ses = [scenes.batch_scene(n, k, seed=i) for i, (n, k) in enumerate([(1200, 0), (900, 1), (700, 2)])]
    pool = cnhip.Solver(0)
    try:
        pool.reserve(sum(sc.n for sc in ses), 64)
        pool.upload_batch(ses)
        for _ in range(3):
            pool.step(params())
        for k in (0, 1, 2):  # reset every kind of scene: PBF, tet block, cloth
            pool.set_scene_state(k, ses[k].pos, ses[k].vel)
            pos, prev, vel = pool.readback_scene(k)
            np.testing.assert_array_equal(pos, ses[k].pos)
            np.testing.assert_array_equal(prev, ses[k].pos)
            np.testing.assert_array_equal(vel, ses[k].vel)
        for _ in range(2):
            pool.step(params())
        for k in (0, 1, 2):
            pos, _, vel = pool.readback_scene(k)
            apos, _, avel = alone(ses[k], 2)
            np.testing.assert_array_equal(pos, apos, err_msg=f"scene {k}")
            np.testing.assert_array_equal(vel, avel, err_msg=f"scene {k}")
    finally:
        pool.close()


def test_reset_only_touches_its_scene_and_takes_tensors():
    torch = pytest.importorskip("torch")
    ses = [scenes.batch_scene(n, k, seed=10 + i) for i, (n, k) in enumerate([(1000, 0), (800, 0)])]
    pool = cnhip.Solver(0)
    try:
        pool.reserve(sum(sc.n for sc in ses), 64)
        pool.upload_batch(ses)
        for _ in range(2):
            pool.step(params())
        p0, _, v0 = pool.readback_scene(0)
        pos = torch.from_numpy(ses[1].pos).cuda()
        vel = torch.from_numpy(ses[1].vel).cuda()
        pool.set_scene_state(1, pos, vel)
        p1, _, v1 = pool.readback_scene(1)
        np.testing.assert_array_equal(p1, ses[1].pos)
        np.testing.assert_array_equal(v1, ses[1].vel)
        q0, _, w0 = pool.readback_scene(0)
        np.testing.assert_array_equal(q0, p0)
        np.testing.assert_array_equal(w0, v0)
        with pytest.raises(ValueError):
            pool.set_scene_state(1, ses[0].pos, ses[0].vel)  # wrong particle count
    finally:
        pool.close()


def test_dirty_runs_equal_whole_scene_writes():
    """cn_particles_update_range: runs of edited particles (one inside a scene, one across a scene
    boundary) give the bits of writing the edited scenes whole, and leave the rest alone."""
    ses = [scenes.batch_scene(n, k, seed=20 + i) for i, (n, k) in enumerate([(1000, 0), (800, 1), (600, 0)])]
    runs = [(37, 5), (ses[0].n - 3, 9)]  # the second run starts in scene 0 and ends in scene 1
    pools = [cnhip.Solver(0), cnhip.Solver(0)]
    try:
        for pool in pools:
            pool.reserve(sum(sc.n for sc in ses), 64)
            pool.upload_batch(ses)
            for _ in range(2):
                pool.step(params())
        pos, _, vel = pools[0].readback()
        for first, count in runs:
            pos[first:first + count, 1] += np.float32(0.004)
            vel[first:first + count, 0] = np.float32(0.25)
            pools[0].update_particles(first, pos[first:first + count], vel[first:first + count])
        for k in (0, 1):  # the scenes the runs touch, written whole
            first, count = pools[1].ranges[k]
            pools[1].set_scene_state(k, pos[first:first + count], vel[first:first + count])
        a0 = pools[0].readback()
        np.testing.assert_array_equal(a0[0], pos)
        np.testing.assert_array_equal(a0[2], vel)
        for pool in pools:
            pool.step(params())
        a, b = pools[0].readback(), pools[1].readback()
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
        with pytest.raises(cnhip.CnError):
            pools[0].update_particles(sum(sc.n for sc in ses) - 2, pos[:3], vel[:3])  # past the end
    finally:
        for pool in pools:
            pool.close()
