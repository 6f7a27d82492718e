"""GPU parity of the PBF substep path against the CPU oracle (run through libcn_hip.so).

Bar: bit-exact keys, sort permutation, cell tables and neighbor lists; positions and velocities
within 1e-5 x particle radius per step from identical state (in practice bitwise equal, which
the tests also report).
"""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

pytestmark = pytest.mark.gpu
F = np.float32


def params(cap=64, substeps=4, iters=4, key_mode=cnhip.CN_KEY_MORTON, **kw):
    return cnhip.default_params(dt=1 / 120, substeps=substeps, default_iterations=iters,
                                max_fluid_neighbors=cap, key_mode=key_mode, **kw)


def mixed_scene(n=3000, seed=9):
    """Fluid in two environments and two materials, solid particles, and a pinned fluid particle:
    exercises the per-class filters of the neighbor walk and the same-material test."""
    rng = np.random.default_rng(seed)
    r = 0.02
    pts = rng.random((n, 3)).astype(np.float32) * F(0.6) + F(0.05)
    sc = scenes._fluid_scene(pts, r, scenes.container((0, 0, 0), (0.7, 0.7, 0.7)), "mixed", {})
    sc.fluid_mats = np.concatenate([cnhip.fluid_material(r), cnhip.fluid_material(r, viscosity=50.0,
                                                                                  cohesion=0.5)])
    sc.fluid_mat = (rng.random(n) < 0.3).astype(np.uint32)
    sc.env = (rng.random(n) < 0.2).astype(np.uint32)
    sc.kinds[rng.random(n) < 0.1] = cnhip.CN_PARTICLE_KIND_SOLID
    sc.pos[7, 3] = 0.0  # pinned fluid particle
    sc.layer[:] = 1
    sc.mask[:] = 0xFFFFFFFF
    sc.mask[rng.random(n) < 0.05] = 2  # cannot see layer-1 particles
    sc.part_mats = np.array([[0.0, 0.0, 0.0, -1.0]], np.float32)
    return sc


def small_cloud(n=1500, seed=5):
    rng = np.random.default_rng(seed)
    return scenes._fluid_scene(rng.random((n, 3)).astype(np.float32) * F(0.5) + F(0.05), 0.02,
                               scenes.container((0, 0, 0), (0.6, 0.6, 0.6)), "cloud", {})


def run_pair(sc, p, steps, solver, oracle, resync=True):
    """Step GPU and oracle from the same state; returns per-step (max |dx|, max |dv|, bitwise)."""
    osc = oracle.OracleScene(sc)
    solver.reserve(sc.n, max(p.max_fluid_neighbors, 64))
    solver.upload(sc)
    out = []
    r = float(np.max(sc.radii))
    for s in range(steps):
        if resync and s > 0:
            cur = sc.copy()
            cur.pos = osc.pos.copy()
            cur.vel = osc.vel.copy()
            solver.upload(cur)
        solver.step(p)
        oracle.step(osc, p)
        pos, prev, vel = solver.readback()
        dx = float(np.abs(pos[:, :3] - osc.pos[:, :3]).max())
        dv = float(np.abs(vel[:, :3] - osc.vel[:, :3]).max())
        bit = bool((pos == osc.pos).all() and (vel == osc.vel).all())
        out.append((dx, dv, bit))
        assert np.isfinite(pos).all() and np.isfinite(vel).all()
        assert dx <= 1e-5 * r, (s, dx, dv)
    m = solver.meta()
    assert m.fluid_neighbor_truncations == 0
    return out


def test_pass_keys_bitexact(solver, oracle):
    rng = np.random.default_rng(1)
    pts = np.zeros((5000, 4), np.float32)
    pts[:, :3] = rng.uniform(-3, 3, (5000, 3))
    pts[:100, :3] = np.round(pts[:100, :3] / 0.1) * F(0.1)  # knife edges at cell multiples
    for cell in (0.1, 0.09999999, 0.0287, 0.25):
        k1, c1 = solver.pass_keys(pts, cell)
        k2, c2 = oracle.broadphase_entries(pts, cell)
        np.testing.assert_array_equal(c1, c2)
        np.testing.assert_array_equal(k1, k2)


@pytest.mark.parametrize("n", [1, 2, 63, 4095, 4096, 4097, 100_003])
def test_pass_sort_bitexact(solver, oracle, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    keys[rng.random(n) < 0.3] = 12345  # many duplicates exercise stability
    idx = np.arange(n, dtype=np.uint32)
    a = solver.pass_sort(keys, idx)
    b = oracle.radix_sort_1bit(keys, idx)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_pass_cell_ranges_bitexact(solver, oracle):
    rng = np.random.default_rng(2)
    for n in (1, 10, 5000, 70_000):
        keys = np.sort(rng.integers(0, max(n // 3, 1), n).astype(np.uint32))
        cap = oracle.next_pow2(max(2 * n, 1))
        np.testing.assert_array_equal(solver.pass_cell_ranges(keys, cap), oracle.cell_ranges(keys, cap))


@pytest.mark.parametrize("builder", ["dam", "cloud", "mixed"])
@pytest.mark.parametrize("key_mode", [cnhip.CN_KEY_HASH, cnhip.CN_KEY_MORTON])
def test_substep_broadphase_and_neighbor_lists_bitexact(solver, oracle, builder, key_mode):
    sc = {"dam": lambda: scenes.dam_break_64k(side=14), "cloud": small_cloud, "mixed": mixed_scene}[builder]()
    p = params(cap=64, key_mode=key_mode)
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    keys, ids, counts, lists = solver.pass_substep_broadphase(p)
    osc = oracle.OracleScene(sc)
    cell = float(np.max(sc.fluid_mats["smoothing_radius"]))
    ref = oracle.substep_neighbors(osc, p, max(cell, 2 * float(sc.radii.max())))
    if key_mode == cnhip.CN_KEY_HASH:
        np.testing.assert_array_equal(keys, ref["sorted_keys"])
        np.testing.assert_array_equal(ids, ref["sorted_idx"])
    else:
        # Morton order: every cell's particles contiguous and in ascending index
        assert (np.diff(keys.astype(np.int64)) >= 0).all()
        same = keys[1:] == keys[:-1]
        assert (ids[1:][same] > ids[:-1][same]).all()
        assert sorted(ids.tolist()) == list(range(sc.n))
    fluid_ok = (sc.kinds == 1) & (sc.pos[:, 3] > 1e-6)
    np.testing.assert_array_equal(counts[fluid_ok], ref["counts"][fluid_ok])
    for i in np.nonzero(fluid_ok)[0]:
        c = counts[i]
        np.testing.assert_array_equal(lists[i, :c], ref["lists"][i, :c])


@pytest.mark.parametrize("builder", ["dam", "cloud", "mixed"])
def test_step_parity_small(solver, oracle, builder):
    sc = {"dam": lambda: scenes.dam_break_64k(side=12), "cloud": small_cloud, "mixed": mixed_scene}[builder]()
    res = run_pair(sc, params(cap=64), 6, solver, oracle, resync=False)
    print(builder, res)
    assert all(b for _, _, b in res), res


def test_step_parity_hash_equals_morton(solver, oracle):
    sc = small_cloud(4000, seed=3)
    out = {}
    for km in (cnhip.CN_KEY_HASH, cnhip.CN_KEY_MORTON):
        solver.reserve(sc.n, 64)
        solver.upload(sc)
        for _ in range(4):
            solver.step(params(cap=64, key_mode=km))
        out[km] = solver.readback()
    for a, b in zip(out[cnhip.CN_KEY_HASH], out[cnhip.CN_KEY_MORTON]):
        np.testing.assert_array_equal(a, b)


def quotient_edge_scene(n=1500, seed=21):
    """A cloud around the origin: coordinates exactly 0, below 2^-62 in magnitude, and pairs a
    hair apart, with a second material that has surface tension and cohesion.  Those particles
    take the IEEE divisions of the neighbor terms; the rest take the reciprocal form."""
    rng = np.random.default_rng(seed)
    r = 0.02
    pts = (rng.random((n, 3)).astype(np.float32) - F(0.5)) * F(0.5)
    pts[0:60, 0] = 0.0
    pts[60:90, 1] = F(1e-25)
    pts[90:110, 2] = F(-3e-22)
    pts[110:120, 0] = F(2.0 ** -70)
    pts[200] = pts[201] + np.array([1e-30, 0, 0], np.float32)
    pts[202] = pts[203]
    sc = scenes._fluid_scene(pts, r, scenes.container((-0.3, -0.3, -0.3), (0.3, 0.3, 0.3)), "qedge", {})
    sc.fluid_mats = np.concatenate([cnhip.fluid_material(r), cnhip.fluid_material(r, cohesion=0.5,
                                                                                  surface_tension=0.2)])
    sc.fluid_mat = (rng.random(n) < 0.4).astype(np.uint32)
    return sc


def huge_radius_scene(n=40, seed=22):
    """Smoothing radius above 1024: the neighbor terms take the IEEE divisions."""
    rng = np.random.default_rng(seed)
    r = 600.0
    sc = scenes._fluid_scene(rng.random((n, 3)).astype(np.float32) * F(3000.0), r,
                             scenes.container((-100, -100, -100), (3100, 3100, 3100), thickness=50.0),
                             "huge", {})
    assert float(sc.fluid_mats["smoothing_radius"][0]) > 1024.0
    return sc


@pytest.mark.parametrize("builder", ["edges", "huge_h"])
def test_step_parity_quotient_edge_cases(solver, oracle, builder):
    sc = {"edges": quotient_edge_scene, "huge_h": huge_radius_scene}[builder]()
    res = run_pair(sc, params(cap=64), 4, solver, oracle, resync=False)
    print(builder, res)
    assert all(b for _, _, b in res), res


def test_step_parity_dam_break_64k(solver, oracle):
    sc = scenes.dam_break_64k()
    res = run_pair(sc, params(cap=64), 3, solver, oracle)
    print(res)
    assert all(b for _, _, b in res), res


def test_step_parity_random_fill_256k(solver, oracle):
    n = 262_144
    sc = scenes.large_pbf(n, box=scenes.scaled_box_for(n))
    res = run_pair(sc, params(cap=64), 1, solver, oracle)
    print(res)
    assert all(b for _, _, b in res), res


def test_large_pbf_4m_properties(solver):
    sc = scenes.large_pbf()
    p = params(cap=64)
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    solver.step(p)
    pos, prev, vel = solver.readback()
    m = solver.meta()
    assert m.fluid_neighbor_truncations == 0 and m.max_fluid_neighbor_count <= 64
    kbar = m.fluid_neighbor_total / (sc.n * p.substeps)
    assert 18 < kbar < 34, kbar
    assert np.isfinite(pos).all() and np.isfinite(vel).all()
    r = sc.radii[0]
    assert (pos[:, :3] >= r - 1e-6).all() and (pos[:, 0] <= 4 - r + 1e-6).all()
    assert (pos[:, 1:3] <= 2 - r + 1e-6).all()
    # the permutation is a bijection: every particle index exported exactly once
    assert np.abs(pos[:, 3] - sc.pos[:, 3]).max() == 0


@pytest.mark.parametrize("case", ["dam_break_64k", "large_pbf_4m"])
def test_candidate_reuse_equals_full_walk(solver, case):
    """Later iterations filter the iteration-0 candidate lists (radius h + skin, the skin 2.5 times
    the previous substep's largest displacement within [0.05 h, 0.3 h]) while every fluid particle
    has moved less than 0.49 skin; the lists, and so the bits, equal the 27-cell walk's."""
    sc = scenes.dam_break_64k() if case == "dam_break_64k" else scenes.large_pbf()
    steps = 6 if case == "dam_break_64k" else 3
    out, reuses = {}, {}
    for flags in (cnhip.abi.CN_STEP_CANDIDATE_REUSE, 0):
        print(f"{case}: flags {flags}", flush=True)
        p = params(cap=64, flags=flags)
        solver.reserve(sc.n, 64)
        solver.upload(sc)
        reuses[flags] = 0
        for _ in range(steps):
            solver.step(p)
            reuses[flags] += solver.meta().candidate_reuses
        out[flags] = solver.readback()
    print(reuses)
    assert reuses[0] == 0 and reuses[cnhip.abi.CN_STEP_CANDIDATE_REUSE] > 0
    for a, b in zip(out[0], out[cnhip.abi.CN_STEP_CANDIDATE_REUSE]):
        np.testing.assert_array_equal(a, b)
