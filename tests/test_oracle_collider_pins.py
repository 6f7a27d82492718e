"""Pins of the oracle's static colliders of any shape (CPU only): ComputeFluidBoundaryContactInfo
for spheres, oriented boxes and capsules against float64 closed forms, the body / local pose
composition and scale rules of physics_shape_common.hlsli, the broad-phase AABB, and a box given
as a shape collider with identity orientation stepping bit for bit like the plain box."""
import math

import numpy as np
import pytest

import cnhip
from cnhip import scenes

SPHERE, BOX, CAPSULE = 0, 1, 2
MARGIN = 0.05


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


def qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return (aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
            aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz)


def box_sdf(local, he):
    """float64 box signed distance and local normal (outside: to the closest point; inside: the
    nearest face, ClosestFaceNormalLocal's tie order x, y, z)."""
    closest = np.clip(local, -he, he)
    d = local - closest
    dl = np.linalg.norm(d)
    if dl > 1e-6:
        return dl, d / dl
    dist = he - np.abs(local)
    n = np.zeros(3)
    if dist[0] <= dist[1] and dist[0] <= dist[2]:
        n[0] = 1.0 if local[0] >= 0 else -1.0
    elif dist[1] <= dist[2]:
        n[1] = 1.0 if local[1] >= 0 else -1.0
    else:
        n[2] = 1.0 if local[2] >= 0 else -1.0
    return -dist.min(), n


def check(oracle, col, p, sd, n, tol=2e-6):
    hit, got_sd, got_n, _ = oracle.collider_contact(col, p)
    assert hit
    assert got_sd == pytest.approx(sd, abs=tol * max(1.0, abs(sd)))
    np.testing.assert_allclose(got_n, n, atol=tol * 4)


def test_sphere_contact_and_scale(oracle):
    # SphereRadius = radius x the largest |scale| component
    col = cnhip.shape_collider(SPHERE, (1.0, 0.5, 0.2), params=(0.3,), scale=(1.0, -2.0, 0.5))
    for p in [(1.9, 0.5, 0.2), (1.1, 0.6, 0.25), (0.2, 1.4, -0.3)]:
        d = np.asarray(p, np.float64) - (1.0, 0.5, 0.2)
        dist = np.linalg.norm(d)
        check(oracle, col, p, dist - 0.6, d / dist)
    # at the centre the normal falls back to +y
    hit, sd, n, _ = oracle.collider_contact(col, (1.0, 0.5, 0.2))
    assert sd == pytest.approx(-0.6) and tuple(n) == (0.0, 1.0, 0.0)


def test_oriented_box_contact(oracle):
    q = quat((0.3, 1.0, -0.2), 0.7)
    R = rotmat(q)
    c = np.array([0.5, 1.0, -0.4])
    he = np.array([0.3, 0.1, 0.2])
    col = cnhip.shape_collider(BOX, c, orientation=q, params=tuple(he))
    rng = np.random.default_rng(5)
    for _ in range(40):
        local = rng.uniform(-0.6, 0.6, 3)
        p = c + R @ local
        sd, nl = box_sdf(local, he)
        if np.min(np.abs(np.abs(local) - he)) < 1e-3:  # skip ties between faces
            continue
        check(oracle, col, p.astype(np.float32), sd, R @ nl, tol=3e-6)


def test_capsule_contact(oracle):
    q = quat((1.0, 0.0, 1.0), 1.1)
    R = rotmat(q)
    c = np.array([0.2, 0.3, 0.4])
    # CapsuleSegment: radius x max(|sx|, |sz|), half height x |sy| along the local y axis
    col = cnhip.shape_collider(CAPSULE, c, orientation=q, params=(0.1, 0.25), scale=(2.0, 1.5, 0.5))
    r, hh = 0.2, 0.375
    axis = R @ np.array([0.0, 1.0, 0.0])
    a, b = c - axis * hh, c + axis * hh
    rng = np.random.default_rng(6)
    for _ in range(40):
        p = c + rng.uniform(-0.8, 0.8, 3)
        t = np.clip(np.dot(p - a, b - a) / np.dot(b - a, b - a), 0.0, 1.0)
        d = p - (a + t * (b - a))
        dist = np.linalg.norm(d)
        check(oracle, col, p.astype(np.float32), dist - r, d / dist, tol=3e-6)


def test_body_and_local_pose_compose(oracle):
    """The contact uses position = body position + R(body) local position and orientation
    normalize(q_body q_local), the inputs normalized first (QuaternionNormalize)."""
    qb, ql = quat((0.0, 1.0, 0.0), 0.5), quat((1.0, 0.0, 0.0), -0.8)
    bp, lp = np.array([1.0, 0.2, 0.0]), np.array([0.3, -0.1, 0.2])
    he = np.array([0.2, 0.15, 0.1])
    col = cnhip.shape_collider(BOX, bp, orientation=tuple(2.0 * np.array(qb)), params=tuple(he), local_position=lp,
                               local_orientation=tuple(0.5 * np.array(ql)))
    cw = bp + rotmat(qb) @ lp
    R = rotmat(qmul(qb, ql))
    rng = np.random.default_rng(7)
    for _ in range(30):
        local = rng.uniform(-0.4, 0.4, 3)
        if np.min(np.abs(np.abs(local) - he)) < 1e-3:
            continue
        sd, nl = box_sdf(local, he)
        check(oracle, col, (cw + R @ local).astype(np.float32), sd, R @ nl, tol=4e-6)


def test_broad_phase_aabb(oracle):
    """ComputeBodyAabb plus the margin: a sphere's cube, a rotated box's projected extents, a
    capsule's segment half-length along |axis| plus its radius; from the scaled local position."""
    col = cnhip.shape_collider(SPHERE, (1.0, 2.0, 3.0), params=(0.5,), scale=(1.0, 3.0, 1.0),
                               local_position=(0.1, 0.0, 0.0))
    _, _, _, aabb = oracle.collider_contact(col, (0.0, 0.0, 0.0))
    np.testing.assert_allclose(aabb, [1.1 - 1.5 - MARGIN, 2 - 1.5 - MARGIN, 3 - 1.5 - MARGIN,
                                      1.1 + 1.5 + MARGIN, 2 + 1.5 + MARGIN, 3 + 1.5 + MARGIN], atol=1e-6)
    q = quat((0.0, 0.0, 1.0), math.pi / 4)
    col = cnhip.shape_collider(BOX, (0.0, 0.0, 0.0), orientation=q, params=(1.0, 0.5, 0.25))
    _, _, _, aabb = oracle.collider_contact(col, (0.0, 0.0, 0.0))
    e = (1.0 + 0.5) / math.sqrt(2.0)
    np.testing.assert_allclose(aabb, [-e - MARGIN, -e - MARGIN, -0.25 - MARGIN, e + MARGIN, e + MARGIN,
                                      0.25 + MARGIN], atol=1e-6)
    col = cnhip.shape_collider(CAPSULE, (0.0, 0.0, 0.0), orientation=quat((1.0, 0.0, 0.0), math.pi / 2),
                               params=(0.1, 0.4))
    _, _, _, aabb = oracle.collider_contact(col, (0.0, 0.0, 0.0))
    np.testing.assert_allclose(aabb, [-0.1 - MARGIN, -0.1 - MARGIN, -0.5 - MARGIN, 0.1 + MARGIN, 0.1 + MARGIN,
                                      0.5 + MARGIN], atol=1e-6)


def test_disabled_collider_has_no_contact(oracle):
    col = cnhip.shape_collider(SPHERE, (0.0, 0.0, 0.0), params=(1.0,), enabled=0)
    assert not oracle.collider_contact(col, (0.5, 0.0, 0.0))[0]


def _as_shape_box(box):
    return cnhip.shape_collider(BOX, tuple(box["center"]), params=tuple(box["half_extents"]),
                                material=tuple(box["material"]), environment=int(box["environment"]),
                                layer=int(box["layer"]), mask=int(box["mask"]), enabled=int(box["enabled"]))


def test_identity_shape_box_steps_like_the_plain_box(oracle):
    sc = scenes.dam_break_64k(side=10)
    alt = sc.copy()
    alt.shape_colliders = np.concatenate([_as_shape_box(b) for b in sc.colliders])
    alt.colliders = sc.colliders[:0]
    p = cnhip.default_params(dt=1 / 120, substeps=2, default_iterations=3, max_fluid_neighbors=64)
    a, b = oracle.OracleScene(sc), oracle.OracleScene(alt)
    for _ in range(3):
        oracle.step(a, p)
        oracle.step(b, p)
    np.testing.assert_array_equal(a.pos, b.pos)
    np.testing.assert_array_equal(a.vel, b.vel)


def test_fluid_rests_outside_a_sphere_and_a_capsule(oracle):
    """A block of fluid dropped on a sphere and a capsule: no particle ends inside either."""
    sc = scenes.shapes_scene(side=10)
    p = cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    osc = oracle.OracleScene(sc)
    for _ in range(30):
        oracle.step(osc, p)
    assert np.isfinite(osc.pos).all()
    r = float(sc.radii[0])
    for col in sc.shape_colliders:
        sds = np.array([oracle.collider_contact(col, q)[1] for q in osc.pos[:, :3]])
        assert sds.min() > 0.5 * r, (int(col["shape"]), sds.min())


def test_out_of_reach_colliders_change_nothing(oracle):
    """Forty pegs far above the fluid, listed after or around the scene's own colliders (49
    colliders in all): the scene steps bit for bit as it does without them, so neither their
    number nor their place in the LBVH visit order reaches the contacts of the colliders in reach."""
    sc = scenes.shapes_scene(side=8)
    far = scenes.pegboard_colliders(8, 5, y=1.6, origin=(2.2, 0.3))
    after, around = sc.copy(), sc.copy()
    after.shape_colliders = np.concatenate([sc.shape_colliders, far])
    around.shape_colliders = np.concatenate([far[:20], sc.shape_colliders, far[20:]])
    p = cnhip.default_params(dt=1 / 120, substeps=2, default_iterations=3, max_fluid_neighbors=64)
    runs = [oracle.OracleScene(x) for x in (sc, after, around)]
    for _ in range(20):
        for o in runs:
            oracle.step(o, p)
    for o in runs[1:]:
        np.testing.assert_array_equal(o.pos, runs[0].pos)
        np.testing.assert_array_equal(o.vel, runs[0].vel)


def test_fluid_and_cloth_rest_on_a_pegboard(oracle):
    """A fluid block and a cloth dropped on 60 pegs (spheres, capsules, turned boxes; 66 colliders
    with the walls): no particle ends inside a peg, and particles rest on them."""
    sc = scenes.pegboard()
    assert len(sc.shape_colliders) + len(sc.colliders) > 32
    p = cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    osc = oracle.OracleScene(sc)
    for _ in range(40):
        oracle.step(osc, p)
    assert np.isfinite(osc.pos).all()
    r = float(sc.radii[0])
    closest = np.full(sc.n, np.inf)
    for col in sc.shape_colliders:
        sds = np.array([oracle.collider_contact(col, q)[1] for q in osc.pos[:, :3]])
        assert sds.min() > 0.5 * r, (int(col["shape"]), sds.min())
        closest = np.minimum(closest, sds)
    assert (closest < 1.1 * r).sum() > 20  # resting on the pegs, not fallen past them
