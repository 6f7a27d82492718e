"""The launch-shape autotuner: tuning leaves the state untouched, and every launch shape (the
tuned one, tiny LDS tiles that send most cell blocks through the run-table walk, a minimal
persistent grid) gives the same bits."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

pytestmark = pytest.mark.gpu


def params(**kw):
    return cnhip.default_params(dt=1 / 120, substeps=2, default_iterations=3, max_fluid_neighbors=64, **kw)


def stepped(sc, p, steps, shape=None, tune=False):
    with cnhip.Solver(0) as s:
        s.reserve(sc.n, 64)
        s.upload(sc)
        s.step(p)
        if shape:
            s.tune_set(*shape)
        if tune:
            before = s.readback()
            r = s.autotune(p, steps=1)
            print("autotune", r, flush=True)
            assert r["candidates"] == 16 and r["ms_per_step"] <= r["default_ms_per_step"]
            after = s.readback()
            for a, b in zip(before, after):
                np.testing.assert_array_equal(a, b)
        for _ in range(steps):
            s.step(p)
        return s.readback()


@pytest.mark.parametrize("shape", [None, (256, 8), (1024, 2048), (3072, 256)])
def test_shapes_give_identical_bits(shape):
    sc = scenes.dam_break_64k(side=24)
    p = params()
    print(f"dam break {sc.n}, shape {shape}", flush=True)
    ref = stepped(sc, p, 3)
    got = stepped(sc, p, 3, shape=shape, tune=shape is None)
    for a, b in zip(ref, got):
        np.testing.assert_array_equal(a, b)


def test_autotune_on_a_pool_keeps_state_and_bits():
    batch = [scenes.batch_scene(n, k, seed=i) for i, (n, k) in enumerate([(3000, 0), (900, 1), (1600, 2), (2500, 0)])]
    p = params()
    with cnhip.Solver(0) as a, cnhip.Solver(0) as b:
        for s in (a, b):
            s.reserve(sum(sc.n for sc in batch), 64)
            s.upload_batch(batch)
            s.step(p)
        r = b.autotune(p, steps=1)
        print("autotune pool", r, flush=True)
        for s in (a, b):
            s.step(p)
            s.step(p)
        for k in range(len(batch)):
            for x, y in zip(a.readback_scene(k), b.readback_scene(k)):
                np.testing.assert_array_equal(x, y)


def test_tune_set_rejects_bad_shapes(solver):
    with pytest.raises(cnhip.CnError):
        solver.tune_set(100, 1024)
    with pytest.raises(cnhip.CnError):
        solver.tune_set(2048, 4)


def test_reserve_rejects_lists_above_4gib():
    # list entries are addressed with 32-bit byte offsets: 2^26 particles x 64 slots x 2 bytes is
    # 8 GiB, refused before anything is allocated
    s = cnhip.Solver(0)
    try:
        with pytest.raises(cnhip.CnError) as e:
            s.reserve(1 << 26, 64)
        assert e.value.status == cnhip.abi.CN_ERR_CAPACITY
    finally:
        s.close()
