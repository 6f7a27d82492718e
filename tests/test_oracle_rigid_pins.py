"""Pins of the oracle's particle-rigid contacts (CPU only): the contact a solid particle generates
against a sphere, an oriented box and a capsule (physics_particle_rigid_generate_contacts.cs.hlsl)
against float64 closed forms, the position solve against a static floor (penetration minus the
slop, capped, times the relaxation), position friction that holds a slow particle and lets a fast
one slide, the restitution-only velocity solve, the candidate filters, two walls summed at a
corner, and soft bodies coming to rest on the floor of their container."""
import math

import numpy as np
import pytest

import cnhip
from cnhip import scenes

SPHERE, BOX, CAPSULE = 0, 1, 2
SLOP, RELAX, MAX_CORR = 1.0e-3, 0.95, 0.05      # solver_config[0], [2], [3] defaults
REST_V, REST_P = 0.5, 2.0                        # solver_config[4], [5] defaults
DT = 1.0 / 120.0
G = -9.81
F = np.float32


def quat(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    s = math.sin(angle / 2)
    return (axis[0] * s, axis[1] * s, axis[2] * s, math.cos(angle / 2))


def rotmat(q):
    x, y, z, w = (float(v) for v in q)
    n = math.sqrt(x * x + y * y + z * z + w * w)
    x, y, z, w = x / n, y / n, z / n, w / n
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def check(oracle, col, p, r, n, pen, tol=3e-6):
    hit, got_n, got_pen = oracle.rigid_contact(col, np.asarray(p, np.float32), r)
    assert hit
    assert got_pen == pytest.approx(pen, abs=tol)
    np.testing.assert_allclose(got_n, n, atol=tol * 4)


def test_sphere_contact(oracle):
    q = quat((0.2, 1.0, 0.4), 0.9)
    c = np.array([0.3, 0.4, -0.2])
    col = cnhip.shape_collider(SPHERE, c, orientation=q, params=(0.25,))
    rng = np.random.default_rng(11)
    for _ in range(30):
        d = rng.normal(size=3)
        p = c + d / np.linalg.norm(d) * rng.uniform(0.05, 0.28)
        dist = np.linalg.norm(p - c)
        check(oracle, col, p, 0.04, (p - c) / dist, 0.04 - (dist - 0.25))
    # at the centre the local normal (0, 1, 0) is rotated into the world
    check(oracle, col, c, 0.04, rotmat(q)[:, 1], 0.29)
    # no contact once the particle's radius clears the surface
    assert not oracle.rigid_contact(col, (c + np.array([0.3, 0.0, 0.0])).astype(np.float32), 0.04)[0]


def test_oriented_box_contact(oracle):
    q = quat((1.0, 0.5, -0.3), 0.6)
    R = rotmat(q)
    c = np.array([1.0, 0.2, 0.5])
    he = np.array([0.3, 0.15, 0.2])
    col = cnhip.shape_collider(BOX, c, orientation=q, params=tuple(he))
    rng = np.random.default_rng(12)
    hits = 0
    for _ in range(60):
        local = rng.uniform(-1.3, 1.3, 3) * (he + 0.02)
        if np.min(np.abs(np.abs(local) - he)) < 2e-3:  # skip ties between faces
            continue
        closest = np.clip(local, -he, he)
        d = local - closest
        if np.linalg.norm(d) > 1e-6:
            sd, nl = np.linalg.norm(d), d / np.linalg.norm(d)
        else:
            depth = he - np.abs(local)
            a = int(np.argmin(depth))
            nl = np.zeros(3)
            nl[a] = 1.0 if local[a] >= 0 else -1.0
            sd = -depth.min()
        r = 0.03
        p = (c + R @ local).astype(np.float32)
        if r - sd <= 1e-5:
            assert not oracle.rigid_contact(col, p, r)[0] or r - sd > -1e-5
            continue
        check(oracle, col, p, r, R @ nl, r - sd)
        hits += 1
    assert hits > 20


def test_capsule_contact(oracle):
    q = quat((0.0, 0.3, 1.0), 1.2)
    R = rotmat(q)
    c = np.array([0.5, 0.5, 0.5])
    col = cnhip.shape_collider(CAPSULE, c, orientation=q, params=(0.08, 0.3))
    axis = R @ np.array([0.0, 1.0, 0.0])
    a, b = c - 0.3 * axis, c + 0.3 * axis
    rng = np.random.default_rng(13)
    for _ in range(40):
        p = c + rng.uniform(-0.35, 0.35, 3)
        t = np.clip(np.dot(p - a, b - a) / np.dot(b - a, b - a), 0.0, 1.0)
        d = p - (a + t * (b - a))
        dist = np.linalg.norm(d)
        if 0.05 - (dist - 0.08) <= 1e-5:
            continue
        check(oracle, col, p, 0.05, d / dist, 0.05 - (dist - 0.08))
    # on the axis the world fallback +x is taken into the collider frame and back
    check(oracle, col, c, 0.05, (1.0, 0.0, 0.0), 0.13)


def floor(material=(0.0, 0.0, 0.0, 0.0), env=0, layer=1, mask=0xFFFFFFFF, enabled=1):
    out = np.zeros(1, cnhip.COLLIDER_DTYPE)
    out[0]["center"] = (0.0, -0.05, 0.0)
    out[0]["half_extents"] = (2.0, 0.05, 2.0)
    out[0]["material"] = material
    out[0]["environment"] = env
    out[0]["layer"] = layer
    out[0]["mask"] = mask
    out[0]["enabled"] = enabled
    return out


def solid(points, r, colliders, vel=None, part_mat=(0.0, 0.0, 0.0, 0.0), w=1.0, shapes=None):
    part = scenes._solid_part(np.asarray(points, np.float32), r, w)
    sc = scenes.compose([part], cnhip.fluid_material(0.03), colliders, "rigid_pin")
    if vel is not None:
        sc.vel[:, :3] = np.asarray(vel, np.float32)
    sc.part_mats = np.array([part_mat], np.float32)
    if shapes is not None:
        sc.shape_colliders = shapes
    return sc


def params(gravity=G, iters=1, substeps=1):
    return cnhip.default_params(dt=DT, substeps=substeps, default_iterations=iters, max_fluid_neighbors=64,
                                gravity=(0.0, gravity, 0.0, 0.0))


def run(oracle, sc, p):
    osc = oracle.OracleScene(sc)
    oracle.step(osc, p)
    return osc


def test_floor_pushes_out_by_the_closed_form(oracle):
    """One substep, one iteration: predict, the floor's correction relax * min(max(pen - slop, 0),
    cap) along +y, the velocity update, then the velocity solve zeroing the approach speed."""
    r, y0, v0 = 0.02, 0.021, -0.5
    osc = run(oracle, solid([(0.1, y0, 0.2)], r, floor(), vel=[(0.0, v0, 0.0)]), params())
    v1 = v0 + G * DT
    ys = F(y0) + v1 * DT
    pen = r - ys
    y = ys + min(max(pen - SLOP, 0.0), MAX_CORR) * RELAX
    assert osc.pos[0, 1] == pytest.approx(y, abs=2e-7)
    assert osc.pos[0, 0] == F(0.1) and osc.pos[0, 2] == F(0.2)
    # (y - y0) / dt < 0 still, so the restitution-free velocity solve stops the approach
    assert (y - y0) / DT < 0
    assert osc.vel[0, 1] == pytest.approx(0.0, abs=1e-6)
    # a deep start: the correction is capped
    osc = run(oracle, solid([(0.1, 0.0, 0.2)], 0.2, floor()), params(gravity=0.0))
    assert osc.pos[0, 1] == pytest.approx(MAX_CORR * RELAX, abs=1e-7)


def test_restitution_of_a_shallow_fast_contact(oracle):
    """The velocity solve gives -e * v_n when the approach is faster than the threshold and the
    contact penetration is at most REST_P x slop; restitution takes the larger of the two."""
    r = 0.02
    y0 = r - 0.0015 + 2.0 * DT  # reaches a penetration of about 1.5e-3 after moving 2 m/s for dt
    sc = solid([(0.0, y0, 0.0)], r, floor(material=(0.0, 0.6, 0.0, 0.0)), vel=[(0.0, -2.0, 0.0)],
               part_mat=(0.0, 0.2, 0.0, 0.0))
    osc = run(oracle, sc, params(gravity=0.0))
    ys = F(y0) - 2.0 * DT
    pen = r - ys
    assert 0 < pen <= REST_P * SLOP
    y = ys + max(pen - SLOP, 0.0) * RELAX
    vn = (y - y0) / DT
    assert -vn > REST_V
    assert osc.vel[0, 1] == pytest.approx(-0.6 * vn, rel=1e-5)


@pytest.mark.parametrize("speed", [0.1, 1.2])
def test_position_friction(oracle, speed):
    """A particle sliding in contact with the floor: the tangential displacement -v dt is undone
    while it stays within the static limit (static friction 1 x the clamped penetration) and
    scaled by kinetic friction x penetration / |displacement| beyond it."""
    r = 0.02
    y0 = r - 0.004
    sc = solid([(0.0, y0, 0.0)], r, floor(material=(0.5, 0.0, 0.0, 1.0)), vel=[(speed, 0.0, 0.0)],
               part_mat=(0.5, 0.0, 0.0, 1.0))
    osc = run(oracle, sc, params(gravity=0.0))
    xs = F(speed * DT)
    pen = min(max(r - F(y0) - SLOP, 0.0), MAX_CORR)
    disp = -float(xs)
    if abs(disp) <= pen:          # static: the whole displacement
        fd = disp
    else:                          # kinetic: friction 0.5 x penetration along it
        fd = disp * min(0.5 * pen / abs(disp), 1.0)
    assert osc.pos[0, 0] == pytest.approx(float(xs) + fd * RELAX, abs=2e-7)
    assert osc.pos[0, 1] == pytest.approx(float(F(y0)) + pen * RELAX, abs=2e-7)


def test_corner_sums_both_walls(oracle):
    r = 0.02
    wall = np.zeros(1, cnhip.COLLIDER_DTYPE)
    wall[0]["center"] = (-0.05, 1.0, 0.0)
    wall[0]["half_extents"] = (0.05, 1.0, 2.0)
    wall[0]["layer"] = 1
    wall[0]["mask"] = 0xFFFFFFFF
    wall[0]["enabled"] = 1
    cols = np.concatenate([floor(), wall])
    osc = run(oracle, solid([(0.015, 0.012, 0.3)], r, cols), params(gravity=0.0))
    assert osc.pos[0, 0] == pytest.approx(0.015 + (r - 0.015 - SLOP) * RELAX, abs=2e-7)
    assert osc.pos[0, 1] == pytest.approx(0.012 + (r - 0.012 - SLOP) * RELAX, abs=2e-7)


@pytest.mark.parametrize("case", ["disabled", "other_env", "layer", "pinned"])
def test_candidate_filters(oracle, case):
    kw = dict(enabled=0) if case == "disabled" else dict(env=1) if case == "other_env" else \
        dict(mask=2) if case == "layer" else {}
    sc = solid([(0.0, 0.01, 0.0)], 0.02, floor(**kw), w=0.0 if case == "pinned" else 1.0)
    osc = run(oracle, sc, params(gravity=0.0))
    assert osc.pos[0, 1] == F(0.01)


def test_shapes_hold_a_cloth(oracle):
    """A cloth dropped over the obstacles(): no particle ends inside a sphere, capsule or box by
    more than the slop, and none below the floor."""
    part = scenes.cloth_part(14, 0.04, 0.02, translation=(0.3, 0.35, 0.35))
    sc = scenes.compose([part], cnhip.fluid_material(0.03), scenes.container((0, 0, 0), (4, 2, 2)), "cloth_shapes")
    sc.shape_colliders = scenes.obstacles()
    osc = oracle.OracleScene(sc)
    p = cnhip.default_params(dt=DT, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    for _ in range(40):
        oracle.step(osc, p)
    assert np.isfinite(osc.pos).all()
    for col in sc.shape_colliders:
        pens = [oracle.rigid_contact(col, q, 0.02)[2] for q in osc.pos[:, :3]]
        assert max(pens) < 0.02, (int(col["shape"]), max(pens))
    assert osc.pos[:, 1].min() > 0.02 - 2 * SLOP


@pytest.mark.parametrize("kind", [1, 2])
def test_soft_bodies_rest_on_the_floor(oracle, kind):
    """A tet block dropped 0.05 m and a cloth dropped 1.2 m onto the floor of their container come
    to rest there after 2 s: lowest particle at radius - slop, the block keeps its height, and the
    velocities have died out (a free fall would be at 20 m/s)."""
    sc = scenes.batch_scene(2500, kind, 0)
    h0 = float(np.ptp(sc.pos[:, 1]))
    osc = oracle.OracleScene(sc)
    p = cnhip.default_params(dt=DT, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    for _ in range(240):
        oracle.step(osc, p)
    r = float(sc.radii[0])
    assert np.isfinite(osc.pos).all()
    assert r - 2 * SLOP < osc.pos[:, 1].min() <= r
    assert float(np.ptp(osc.pos[:, 1])) >= 0.95 * h0
    assert np.abs(osc.vel[:, :3]).max() < 0.02
