"""The soft phase beside the fluid passes (cn_concurrent_phases): a pool of fluid, soft-body and
cloth scenes (no particle-particle pairs, so the two chains are independent within a substep) and
the mixed scene steps bit for bit alike with the soft phase on its own stream and on the context
stream, eager and graph-captured."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

pytestmark = pytest.mark.gpu


def params():
    return cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)


def run(batch, concurrent, graphs, steps=3):
    p = params()
    with cnhip.Solver(0) as pool:
        pool.reserve(sum(sc.n for sc in batch), 64)
        pool.upload_batch(batch) if len(batch) > 1 else pool.upload(batch[0])
        pool.concurrent_phases(concurrent)
        pool.graphs(graphs)
        for _ in range(steps):
            pool.step(p)
        return pool.readback()


@pytest.mark.parametrize("graphs", [False, True])
def test_concurrent_phases_equal_serial_on_a_pool(graphs):
    batch = [scenes.batch_scene(2500, 0, 1), scenes.batch_scene(2000, 1, 2), scenes.batch_scene(1800, 2, 3),
             scenes.batch_scene(3500, 0, 4), scenes.batch_scene(1500, 1, 5)]
    print(f"pool of {len(batch)} scenes, graphs {graphs}", flush=True)
    a = run(batch, True, graphs)
    b = run(batch, False, graphs)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_concurrent_phases_equal_serial_on_the_mixed_scene():
    sc = scenes.mixed(24, 10, 48)
    a = run([sc], True, False, steps=4)
    b = run([sc], False, False, steps=4)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
