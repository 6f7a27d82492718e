"""The batched and served workloads at their full sizes: 512 scenes (log-uniform 2k-64k particles,
40 % PBF / 30 % tet / 30 % cloth, seed 3) stepped as one graph-captured pool, and 2,048 sessions of
the same distribution served under a 200 GB budget with Poisson requests at 2,000 req/s.  Every
scene is checked through size-independent properties (finite, inside its container, no neighbor
truncation); a seeded sample of every kind is checked bit for bit against stepping the scene
alone, and the smallest ones against the CPU oracle."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes, serving

pytestmark = pytest.mark.gpu

BOX_LO = np.array([0.0, 0.0, 0.0], np.float32)
BOX_HI = np.array([4.0, 2.0, 2.0], np.float32)


def params():
    return cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)


def sample_per_kind(kinds, per_kind, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(3):
        ids = np.flatnonzero(np.asarray(kinds) == k)
        out += sorted(rng.choice(ids, min(per_kind, len(ids)), replace=False).tolist())
    return out


def check_scene_state(pos, vel, name, slack=0.05):
    assert np.isfinite(pos).all() and np.isfinite(vel).all(), name
    lo = pos[:, :3].min(0)
    hi = pos[:, :3].max(0)
    assert (lo >= BOX_LO - slack).all() and (hi <= BOX_HI + slack).all(), (name, lo, hi)


def step_alone(solver, sc, steps, p):
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    for _ in range(steps):
        solver.step(p)
    return solver.readback()


def test_batched_512_scenes(solver, oracle):
    batch = scenes.batched(512, workers=8)
    _, kinds = scenes.batch_sizes(512)
    p = params()
    steps = 2
    print(f"batched: 512 scenes, {sum(sc.n for sc in batch)} particles, {steps} graph-captured steps", flush=True)
    with cnhip.Solver(0) as pool:
        pool.reserve(sum(sc.n for sc in batch), 64)
        pool.upload_batch(batch)
        pool.graphs(True)
        for _ in range(steps):
            pool.step(p)
        m = pool.meta(check=True)
        assert m.fluid_neighbor_truncations == 0 and m.max_fluid_neighbor_count <= 64
        states = [pool.readback_scene(k) for k in range(len(batch))]
    for k, (pos, _, vel) in enumerate(states):
        check_scene_state(pos, vel, batch[k].name)
        assert not np.array_equal(pos[:, :3], batch[k].pos[:, :3]), f"scene {k} did not move"
    sample = sample_per_kind(kinds, 6, seed=31)
    for k in sample:
        pos, prev, vel = step_alone(solver, batch[k], steps, p)
        np.testing.assert_array_equal(states[k][0], pos, err_msg=f"scene {k} {batch[k].name}")
        np.testing.assert_array_equal(states[k][1], prev, err_msg=f"scene {k} prev")
        np.testing.assert_array_equal(states[k][2], vel, err_msg=f"scene {k} velocities")
    # the four smallest scenes against the CPU oracle
    for k in np.argsort([sc.n for sc in batch])[:4]:
        osc = oracle.OracleScene(batch[k])
        for _ in range(steps):
            oracle.step(osc, p)
        np.testing.assert_array_equal(states[k][0], osc.pos, err_msg=f"scene {k} vs oracle")
        np.testing.assert_array_equal(states[k][2], osc.vel, err_msg=f"scene {k} vs oracle")


def test_served_2048_sessions(solver):
    ses = scenes.batched(2048, workers=8)
    _, kinds = scenes.batch_sizes(2048)
    p = params()
    srv = serving.Server(ses, p, budget_bytes=200 << 30, shards=8, graphs=True)
    try:
        assert len(srv.deferred) == 0 and len(srv.admitted) == 2048
        assert srv.bytes_used <= 200 << 30
        arr = serving.poisson_arrivals(2000.0, 0.5, srv.admitted, seed=4)
        print(f"served: 2048 sessions, {sum(sc.n for sc in ses)} particles, {len(arr)} requests", flush=True)
        done = srv.serve(arr)
        assert len(done) == len(arr)
        assert all(r.t_done >= r.t_arrival for r in done)
        served = np.bincount([r.session for r in done], minlength=len(ses))
        sample = sample_per_kind(kinds, 4, seed=41)
        sample += [int(k) for k in np.flatnonzero(served >= 2)[:4]]
        states = {k: srv.session_state(k) for k in sample}
        for j, sh in enumerate(srv.shards):
            pos, _, vel = sh.solver.readback()
            check_scene_state(pos, vel, f"shard {j}")
            assert sh.solver.meta().fluid_neighbor_truncations == 0
    finally:
        srv.close()
    for k in sample:
        pos, prev, vel = step_alone(solver, ses[k], int(served[k]), p)
        np.testing.assert_array_equal(states[k][0], pos, err_msg=f"session {k} ({served[k]} steps)")
        np.testing.assert_array_equal(states[k][2], vel, err_msg=f"session {k}")
