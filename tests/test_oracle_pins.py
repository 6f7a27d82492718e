"""Pins of the CPU oracle against facts fixed outside it (CPU only).

Each check is chosen so a plausible mistake in the restatement fails it: the reference's own test
assertions (tests/physics/*.cpp), closed forms, brute force on small inputs, and an independent
restatement of the static LBVH traversal.
"""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

F = np.float32


def test_fluid_material_matches_reference_test_facts(oracle):
    # tests/physics/physics_fluid_unified_particle_pool.cpp:60 -- smoothingRadius == (2/0.6) r
    for r in (0.03, 0.0086, 0.2, 0.05):
        m = oracle.fluid_material(r)
        h = max(F(F(2.0) / F(0.6)) * F(r), max(F(2.0) * F(r), F(1e-4)))
        assert m[2] == h
        # invRestDensity is the float reciprocal of the derived rest density
        assert m[1] == F(1.0) / m[0]
        # cohesion1/cohesion2 from the clamped rest ratio (physics_world.cpp:243-246)
        rr = min(max(F(max(F(2) * F(r), F(1e-4))) / h, F(1e-3)), F(0.999))
        assert m[6] == -(F(1) + rr) / (rr * rr)


def test_rest_density_is_the_tight_pack_kernel_sum(oracle):
    # independent float64 restatement of tightPack3D + fluidSpikyWeight (physics_world.cpp:99-231)
    r = 0.03
    m = oracle.fluid_material(r)
    d, h = 2 * r, float(m[2])
    dim = int(np.ceil(h / d))
    row = np.sqrt(0.75) * d
    rho = 0.0
    dcs = 0.0
    for z in range(-dim, dim + 1):
        for y in range(-dim, dim + 1):
            for x in range(-dim, dim + 1):
                s = np.array([x * d + (d * 0.5 if (y + z) & 1 else 0.0), y * row, z * row])
                dist = np.linalg.norm(s)
                if dist == 0 or dist > h:
                    continue
                q = 1 - dist / h
                k = 15 / (np.pi * h ** 3)
                rho += k * q * q
                dcs += (-2 * k * q / h) ** 2
    assert m[0] == pytest.approx(rho, rel=1e-5)
    assert m[3] == pytest.approx(1.0 / dcs, rel=1e-5)


def test_spiky_kernel_integrates_to_two(oracle):
    # 15/(pi h^3) int_0^h (1 - r/h)^2 4 pi r^2 dr = 2: on a fine cubic lattice of spacing s the
    # density of an interior particle approaches 2 / s^3 minus its own W(0) = 15 / (pi h^3)
    s, h = 0.01, 0.05
    g = np.arange(-9, 10, dtype=np.float32) * F(s) + F(1.0)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    sc = scenes._fluid_scene(pts, 0.003, np.zeros(0, cnhip.COLLIDER_DTYPE), "fine", {})
    sc.fluid_mats["smoothing_radius"] = h
    sc.fluid_mats["rest_density"] = 1.0
    sc.fluid_mats["density_constraint_scale"] = 1.0
    p = cnhip.default_params(dt=1 / 120, substeps=1, default_iterations=1, max_fluid_neighbors=1024,
                             gravity=(0.0, 0.0, 0.0, 0.0))
    osc = oracle.OracleScene(sc)
    nb = oracle.substep_neighbors(osc, p, h)
    nc = oracle.fluid_density(osc, p, nb["counts"], nb["lists"])
    centre = int(np.argmin(np.linalg.norm(pts - F(1.0), axis=1)))
    density = float(nc[centre, 3]) + 1.0
    want = 2.0 / s ** 3 - 15 / (np.pi * h ** 3)
    assert density == pytest.approx(want, rel=0.02)


def test_grid_cell_size_reference_facts(oracle):
    # tests/physics/physics_cloth_authoring.cpp:62 -- radius 0.2 solid particles give cell 0.4
    import ctypes as C
    lib = oracle.load()
    radii = np.full(4, 0.2, np.float32)
    assert lib.cno_grid_cell_size(4, radii.ctypes.data, None, 0, None) == F(0.4)
    # with a fluid material the cell is max(2r, h)
    fm = np.zeros(1, cnhip.FLUID_MATERIAL_DTYPE)
    fm["smoothing_radius"] = 0.5
    idx = np.zeros(4, np.uint32)
    assert lib.cno_grid_cell_size(4, radii.ctypes.data, idx.ctypes.data, 1, fm.ctypes.data) == F(0.5)
    assert lib.cno_grid_cell_size(0, radii.ctypes.data, None, 0, None) == F(1e-4)
    del C


def _hash(gx, gy, gz):
    m = 0xFFFFFFFF
    return (((gx * 73856093) & m) ^ ((gy * 19349663) & m) ^ ((gz * 83492791) & m)) & m


def test_cell_keys_floor_and_hash(oracle):
    cell = 0.25  # a power of two: p * (1 / cell) is exact, so floor is decided exactly
    pts = np.array([[0.0, 0.0, 0.0, 1], [-0.25, 0.25, 0.5, 1], [0.2499999, -1e-9, -0.2500001, 1],
                    [1.0, -1.0, 3.75, 1], [-3.0, 7.5, -0.125, 1]], np.float32)
    keys, cells = oracle.broadphase_entries(pts, cell)
    want = np.array([[0, 0, 0], [-1, 1, 2], [0, -1, -2], [4, -4, 15], [-12, 30, -1]], np.int32)
    np.testing.assert_array_equal(cells, want)  # floor, not truncation toward zero
    for k, c in zip(keys, want):
        assert int(k) == _hash(int(c[0]), int(c[1]), int(c[2]))
    assert _hash(1, 0, 0) == 73856093 and _hash(0, 0, 0) == 0


def test_one_bit_radix_sort_is_a_stable_sort(oracle):
    rng = np.random.default_rng(7)
    for n, hi in ((1, 5), (2, 2), (1000, 17), (5000, 1 << 32), (4097, 3)):
        keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
        idx = np.arange(n, dtype=np.uint32)
        sk, si = oracle.radix_sort_1bit(keys, idx)
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(si, order.astype(np.uint32))
        np.testing.assert_array_equal(sk, keys[order])


def test_cell_range_table_invariants(oracle):
    rng = np.random.default_rng(3)
    keys = np.sort(rng.integers(0, 50, 300).astype(np.uint32))
    keys[-4:] = 0xFFFFFFFF  # a real cell whose hash equals the empty marker
    cap = oracle.next_pow2(2 * len(keys))
    assert cap == 1024
    tbl = oracle.cell_ranges(keys, cap)
    uk, first, cnt = np.unique(keys, return_index=True, return_counts=True)
    nr = len(uk)
    np.testing.assert_array_equal(tbl[:nr, 0], uk)
    np.testing.assert_array_equal(tbl[:nr, 1], first)
    np.testing.assert_array_equal(tbl[:nr, 2], first + cnt)
    assert (tbl[nr:, 0] == 0xFFFFFFFF).all() and (tbl[nr:, 1:] == 0).all()
    assert (np.diff(tbl[:nr, 0].astype(np.int64)) > 0).all()
    for k, f, c in zip(uk[:-1], first, cnt):
        assert oracle.find_cell_range(tbl, int(k)) == (int(f), int(f + c))
    assert oracle.find_cell_range(tbl, 51) is None
    # the cell whose key is 0xffffffff is indistinguishable from an empty slot: never found
    assert oracle.find_cell_range(tbl, 0xFFFFFFFF) is None
    assert oracle.next_pow2(0) == 1 and oracle.next_pow2(1) == 1 and oracle.next_pow2(1025) == 2048


def _small_cloud(n=1200, seed=5, box=0.55, r=0.02):
    rng = np.random.default_rng(seed)
    sc = scenes._fluid_scene(rng.random((n, 3)).astype(np.float32) * F(box) + F(0.2), r,
                             np.zeros(0, cnhip.COLLIDER_DTYPE), "cloud", {})
    return sc


def test_neighbor_lists_match_brute_force(oracle):
    sc = _small_cloud()
    h = float(sc.fluid_mats["smoothing_radius"][0])
    cell = h
    p = cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64,
                             gravity=(0.0, -9.81, 0.0, 0.0))
    osc = oracle.OracleScene(sc)
    out = oracle.substep_neighbors(osc, p, cell)
    pos = out["pos"][:, :3]
    cells = np.floor(pos * (F(1) / F(cell))).astype(np.int64)
    hf = F(h)
    for i in range(sc.n):
        want = []
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    t = cells[i] + np.array([dx, dy, dz])
                    js = np.nonzero((cells == t).all(1))[0]
                    for j in js:  # ascending particle index inside a cell
                        if j == i:
                            continue
                        d = pos[i] - pos[j]
                        dist = np.sqrt(F(F(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]))
                        if dist < hf:
                            want.append(j)
        want = want[:64]
        assert out["counts"][i] == len(want)
        np.testing.assert_array_equal(out["lists"][i, :len(want)], np.array(want, np.uint32))


def test_neighbor_cap_truncates_in_scan_order(oracle):
    sc = _small_cloud(n=600, box=0.25)
    p = cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=8)
    osc = oracle.OracleScene(sc)
    full = oracle.substep_neighbors(osc, cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4,
                                                              max_fluid_neighbors=128),
                                    float(sc.fluid_mats["smoothing_radius"][0]))
    cut = oracle.substep_neighbors(osc, p, float(sc.fluid_mats["smoothing_radius"][0]))
    assert (full["required"] > 8).any()
    np.testing.assert_array_equal(cut["counts"], np.minimum(full["counts"], 8))
    np.testing.assert_array_equal(cut["lists"][:, :8], full["lists"][:, :8] * (np.arange(8) < cut["counts"][:, None]))


def _karras_visit_order(codes):
    """Independent restatement of physics_rigid_bvh_hierarchy.cs.hlsl + a right-first stack walk."""
    n = len(codes)
    order = np.argsort(codes, kind="stable")
    sc = [int(codes[k]) for k in order]
    el = [int(k) for k in order]
    if n == 1:
        return [el[0]]

    def delta(i, j):
        if j < 0 or j >= n:
            return -1
        if sc[i] == sc[j]:
            x = el[i] ^ el[j]
            return 64 if x == 0 else 63 - (x.bit_length() - 1)
        return 31 - ((sc[i] ^ sc[j]).bit_length() - 1)

    nodes = {}
    leaf = n - 1
    for idx in range(n - 1):
        dl, dr = delta(idx, idx - 1), delta(idx, idx + 1)
        d = 1 if dr >= dl else -1
        dmin = min(dl, dr)
        lmax = 2
        while delta(idx, idx + lmax * d) > dmin:
            lmax <<= 1
        ln = 0
        s = lmax >> 1
        while s > 0:
            if delta(idx, idx + (ln + s) * d) > dmin:
                ln += s
            s >>= 1
        first, last = min(idx, idx + ln * d), max(idx, idx + ln * d)
        cp = delta(first, last)
        split, stride = first, last - first
        while True:
            stride = (stride + 1) >> 1
            ns = split + stride
            if ns < last and delta(first, ns) > cp:
                split = ns
            if stride <= 1:
                break
        a = leaf + split if split == first else split
        b = leaf + split + 1 if split + 1 == last else split + 1
        nodes[idx] = (a, b)
    visit, stack = [], [0]
    while stack:
        k = stack.pop()
        if k >= leaf:
            visit.append(el[k - leaf])
            continue
        a, b = nodes[k]
        stack.append(a)
        stack.append(b)
    return visit


def _morton(c, lo, hi):
    def expand(v):
        v = (v * 0x00010001) & 0xFF0000FF
        v = (v * 0x00000101) & 0x0F00F00F
        v = (v * 0x00000011) & 0xC30C30C3
        v = (v * 0x00000005) & 0x49249249
        return v
    q = []
    for a in range(3):
        size = max(F(hi[a] - lo[a]), F(1e-5))
        m = min(max(F((c[a] - lo[a]) / size), F(0)), F(1))
        q.append(int(min(max(m * F(1024), F(0)), F(1023))))
    return expand(q[0]) * 4 + expand(q[1]) * 2 + expand(q[2])


def test_static_collider_visit_order_is_the_lbvh_walk(oracle):
    rng = np.random.default_rng(11)
    for trial in range(6):
        nc = [1, 2, 5, 6, 9, 17][trial]
        cols = np.zeros(nc, cnhip.COLLIDER_DTYPE)
        cols["center"] = rng.uniform(-2, 2, (nc, 3))
        if trial == 4:
            cols["center"][3] = cols["center"][1]  # duplicate Morton codes: tie broken by index
        cols["half_extents"] = rng.uniform(0.05, 1.0, (nc, 3))
        cols["enabled"] = 1
        mn = (cols["center"] - cols["half_extents"]).astype(np.float32) - F(0.05)
        mx = (cols["center"] + cols["half_extents"]).astype(np.float32) + F(0.05)
        lo, hi = mn.min(0), mx.max(0)
        codes = np.array([_morton(F(0.5) * (mn[k] + mx[k]), lo, hi) for k in range(nc)], np.uint64)
        want = _karras_visit_order(codes)
        got = oracle.static_collider_order(cols)
        assert list(got) == want
    order = oracle.static_collider_order(scenes.container((0, 0, 0), (4, 2, 2)))
    assert sorted(order.tolist()) == list(range(6))


def test_free_fall_closed_form(oracle):
    # symplectic Euler under gravity: v_k = v_{k-1} + g dt, x_k = x_{k-1} + v_k dt; the reference
    # re-derives v = (x - x_prev) / dt, which loses about half an ulp of x per substep, so the
    # particle starts near the origin where that ulp is small
    sc = scenes._fluid_scene(np.array([[0.125, 0.25, 0.125]], np.float32), 0.03,
                             np.zeros(0, cnhip.COLLIDER_DTYPE), "one", {})
    p = cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    osc = oracle.OracleScene(sc)
    dt = (1 / 120) / 4
    x, v = 0.25, 0.0
    for s in range(10):
        oracle.step(osc, p)
        for k in range(4):
            v += -9.81 * dt
            x += v * dt
        assert osc.vel[0, 1] == pytest.approx(v, abs=3e-5 * (s + 1))
        assert osc.pos[0, 1] == pytest.approx(x, abs=1e-7 * (s + 1) ** 2)
        assert osc.pos[0, 0] == F(0.125) and osc.vel[0, 0] == 0


def test_rest_lattice_has_rest_density(oracle):
    # interior particles of the tight-pack lattice see exactly the lattice the rest density is
    # derived from, so their density equals restDensity (C = 0) up to float summation order
    r = 0.03
    d = 2 * r
    row = np.sqrt(0.75) * d
    pts = []
    for z in range(-7, 8):
        for y in range(-7, 8):
            for x in range(-7, 8):
                pts.append([x * d + (d * 0.5 if (y + z) & 1 else 0.0), y * row, z * row])
    pts = np.array(pts, np.float32) + F(1.0)
    sc = scenes._fluid_scene(pts, r, np.zeros(0, cnhip.COLLIDER_DTYPE), "lattice", {})
    p = cnhip.default_params(dt=1 / 120, substeps=1, default_iterations=1, max_fluid_neighbors=128,
                             gravity=(0.0, 0.0, 0.0, 0.0))
    osc = oracle.OracleScene(sc)
    h = float(sc.fluid_mats["smoothing_radius"][0])
    nb = oracle.substep_neighbors(osc, p, h)
    nc = oracle.fluid_density(osc, p, nb["counts"], nb["lists"])
    centre = np.linalg.norm(pts - F(1.0), axis=1) < 3 * d
    rho0 = float(sc.fluid_mats["rest_density"][0])
    dcs = float(sc.fluid_mats["density_constraint_scale"][0])
    density = nc[centre, 3] / dcs + rho0
    assert np.abs(density / rho0 - 1).max() < 1e-4
    # surface particles are under-dense and clamp at C = max(rho - rho0, 0) * scale = 0
    assert (nc[:, 3] >= 0).all()


def test_clamp_pushes_to_radius_and_projection_removes_inward_velocity(oracle):
    cols = scenes.container((0, 0, 0), (1, 1, 1))
    pts = np.array([[0.5, 0.01, 0.5]], np.float32)
    sc = scenes._fluid_scene(pts, 0.03, cols, "one", {})
    sc.vel[0, 1] = -1.0
    p = cnhip.default_params(dt=1 / 120, substeps=1, default_iterations=2, max_fluid_neighbors=64,
                             gravity=(0.0, 0.0, 0.0, 0.0))
    osc = oracle.OracleScene(sc)
    oracle.step(osc, p)
    assert osc.pos[0, 1] == pytest.approx(0.03, abs=1e-7)
    assert osc.vel[0, 1] >= 0.0
