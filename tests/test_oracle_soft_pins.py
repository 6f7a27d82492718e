"""Pins of the oracle's XPBD soft constraints and of the soft/cloth authoring (CPU only)."""
import math

import numpy as np
import pytest

import cnhip
from cnhip import authoring as au
from cnhip import scenes

F = np.float32


def test_tet_grid_reference_counts():
    # tests/physics/physics_soft_body_generation_regular_tet_grid.cpp:36 -- a 2 m cube at spacing 1
    pos, tets = au.tet_grid((2.0, 2.0, 2.0), spacing=1.0)
    assert len(pos) == 27 and len(tets) == 40
    # 5 tets per cube fill the cube exactly: the rest volumes sum to the cube volume
    _, tc = au.tet_constraints(pos, tets)
    assert float(tc["rest_volume"].astype(np.float64).sum()) == pytest.approx(8.0, rel=1e-6)
    pos, tets = au.tet_grid((0.62, 0.62, 0.62), resolution=(32, 32, 32))
    edges, tc = au.tet_constraints(pos, tets)
    assert len(tets) == 148_955 and len(edges) == 187_488


def test_cloth_reference_counts():
    # tests/physics/physics_cloth_authoring.cpp:21-30 -- a two-triangle quad: 5 edges, 1 dihedral,
    # rest angle 0
    quad = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [1, 0, 1]], np.float32)
    e, d = au.cloth_constraints(quad, np.array([[0, 1, 3], [0, 3, 2]], np.uint32))
    assert len(e) == 5 and len(d) == 1 and d["rest_angle"][0] == 0
    assert list(zip(e["a"], e["b"])) == sorted(zip(e["a"], e["b"]))
    p, tris = au.cloth_grid(256, 256, 0.0075)
    e, d = au.cloth_constraints(p, tris)
    assert len(tris) == 130_050 and len(e) == 195_585 and len(d) == 194_565


def test_cephes_atan2_acos_accuracy(oracle):
    lib = oracle.load()
    import ctypes as C
    lib.cno_atan2f.restype = C.c_float
    lib.cno_atan2f.argtypes = [C.c_float, C.c_float]
    lib.cno_acosf.restype = C.c_float
    lib.cno_acosf.argtypes = [C.c_float]
    rng = np.random.default_rng(0)
    for y, x in list(rng.normal(size=(2000, 2)).astype(np.float32)) + [(0.0, 1.0), (0.0, -1.0), (1.0, 0.0),
                                                                           (-1.0, 0.0), (1e-30, 1.0)]:
        got = lib.cno_atan2f(float(y), float(x))
        want = math.atan2(float(y), float(x))
        assert abs(got - want) <= 3 * np.spacing(F(abs(want)) + F(1e-30)) + 1e-7, (y, x, got, want)
    for x in np.linspace(-0.9999, 0.9999, 4001).astype(np.float32):
        got = lib.cno_acosf(float(x))
        want = math.acos(float(x))
        assert abs(got - want) <= 3 * np.spacing(F(want)) + 1e-7, (x, got, want)


def _solid_scene(pos, inv_mass, edges=None, bends=None, dihedrals=None, tets=None):
    part = scenes._solid_part(np.asarray(pos, np.float32), 0.01, 1.0)
    part["pos"][:, 3] = inv_mass
    for k, v in dict(edges=edges, bends=bends, dihedrals=dihedrals, tets=tets).items():
        if v is not None:
            part[k] = v
    return scenes.compose([part], cnhip.fluid_material(0.03), np.zeros(0, cnhip.COLLIDER_DTYPE), "t")


def _params(iters=1, substeps=1):
    return cnhip.default_params(dt=1 / 120, substeps=substeps, default_iterations=iters, max_fluid_neighbors=64,
                                gravity=(0.0, 0.0, 0.0, 0.0))


def test_edge_closed_form(oracle):
    # two unit-mass particles 2 m apart, rest length 1, compliance 0: one Jacobi solve moves each by
    # relaxation * C / 2 along the edge, so the new length is 2 - 0.2 * 1
    e = np.zeros(1, au.EDGE_DTYPE)
    e["a"], e["b"], e["rest_length"], e["enabled"] = 0, 1, 1.0, 1
    sc = _solid_scene([[0, 0, 0], [2, 0, 0]], 1.0, edges=e)
    osc = oracle.OracleScene(sc)
    oracle.step(osc, _params())
    assert osc.pos[1, 0] - osc.pos[0, 0] == pytest.approx(1.8, abs=1e-6)
    assert osc.pos[0, 0] == pytest.approx(0.1, abs=1e-6)
    # a pinned end takes no correction and the free end takes all of it
    sc.pos[0, 3] = 0.0
    osc = oracle.OracleScene(sc)
    oracle.step(osc, _params())
    assert osc.pos[0, 0] == 0.0 and osc.pos[1, 0] == pytest.approx(1.8, abs=1e-6)


def _tet_volume(p):
    return float(np.dot(np.cross(p[1] - p[0], p[2] - p[0]), p[3] - p[0]) / 6.0)


def test_tet_gradient_matches_finite_differences_and_restores_volume(oracle):
    # the unit right tet has volume 1/6; compress it and the solve must move along -C grad V,
    # and a rigid rotation of a rest tet has C = 0 (no correction)
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
    assert _tet_volume(p) == pytest.approx(1 / 6)
    t = np.zeros(1, au.TET_DTYPE)
    t["p"] = [0, 1, 2, 3]
    t["rest_volume"] = 1 / 6
    q = p.copy()
    q[3, 2] = 0.6
    sc = _solid_scene(q, 1.0, tets=t)
    osc = oracle.OracleScene(sc)
    oracle.step(osc, _params())
    moved = osc.pos[:, :3].astype(np.float64) - q
    grad = np.zeros((4, 3))
    for k in range(4):
        for a in range(3):
            d = np.zeros((4, 3))
            d[k, a] = 1e-6
            grad[k, a] = (_tet_volume(q + d) - _tet_volume(q - d)) / 2e-6
    C = _tet_volume(q) - 1 / 6
    dl = -C / float((grad * grad).sum())
    np.testing.assert_allclose(moved, 0.2 * dl * grad, atol=2e-6)
    assert _tet_volume(osc.pos[:, :3].astype(np.float64)) > _tet_volume(q)
    # rotated rest tet: no motion
    c, s = math.cos(0.7), math.sin(0.7)
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    r = (p @ R.T + 0.5).astype(np.float32)
    t["rest_volume"] = abs(_tet_volume(r.astype(np.float64)))
    sc = _solid_scene(r, 1.0, tets=t)
    osc = oracle.OracleScene(sc)
    oracle.step(osc, _params())
    np.testing.assert_allclose(osc.pos[:, :3], r, atol=1e-6)


def test_bend_and_dihedral_move_toward_rest(oracle):
    b = np.zeros(1, au.BEND_DTYPE)
    b["p0"], b["p1"], b["p2"], b["rest_angle"] = 0, 1, 2, math.pi
    pts = np.array([[-1, 0.3, 0], [0, 0, 0], [1, 0.3, 0]], np.float32)
    sc = _solid_scene(pts, 1.0, bends=b)
    osc = oracle.OracleScene(sc)

    def angle(p):
        u, v = p[0] - p[1], p[2] - p[1]
        return math.acos(np.dot(u, v) / np.linalg.norm(u) / np.linalg.norm(v))
    a0 = angle(pts.astype(np.float64))
    oracle.step(osc, _params(iters=4))
    assert angle(osc.pos[:, :3].astype(np.float64)) > a0
    # a flat hinge has dihedral 0: nothing moves; a folded one opens toward 0
    quad = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [1, 0, 1]], np.float32)
    e, d = au.cloth_constraints(quad, np.array([[0, 1, 3], [0, 3, 2]], np.uint32))
    sc = _solid_scene(quad, 1.0, dihedrals=d)
    osc = oracle.OracleScene(sc)
    oracle.step(osc, _params(iters=4))
    np.testing.assert_array_equal(osc.pos[:, :3], quad)
    fold = quad.copy()
    fold[2, 1] = 0.5  # lift one opposite vertex

    def dihedral(p):
        e0, e1, o0, o1 = p[d["e0"][0]], p[d["e1"][0]], p[d["o0"][0]], p[d["o1"][0]]
        n0 = np.cross(e0 - o0, e1 - o0)
        n1 = np.cross(e1 - o1, e0 - o1)
        return abs(math.atan2(np.dot((e1 - e0) / np.linalg.norm(e1 - e0), np.cross(n0, n1)), np.dot(n0, n1)))
    sc = _solid_scene(fold, 1.0, dihedrals=d)
    osc = oracle.OracleScene(sc)
    oracle.step(osc, _params(iters=4))
    assert dihedral(osc.pos[:, :3].astype(np.float64)) < dihedral(fold.astype(np.float64))


def test_soft_block_holds_volume_under_gravity(oracle):
    sc = scenes.compose([scenes.soft_block_part(5, 0.1, compliance=0.0)], cnhip.fluid_material(0.03),
                        np.zeros(0, cnhip.COLLIDER_DTYPE), "block")
    v0 = float(sc.tets["rest_volume"].astype(np.float64).sum())
    osc = oracle.OracleScene(sc)
    p = cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=8, max_fluid_neighbors=64)
    for _ in range(20):
        oracle.step(osc, p)
    pts = osc.pos[:, :3].astype(np.float64)
    vol = sum(abs(_tet_volume(pts[t])) for t in sc.tets["p"])
    assert abs(vol - v0) / v0 < 0.01
    # free fall of the whole block: centre of mass follows g t^2 / 2 within float noise
    t = 20 / 120
    assert pts[:, 1].mean() == pytest.approx(float(sc.pos[:, 1].mean()) - 0.5 * 9.81 * t * t, abs=2e-3)


# ---- XPBD magnitudes of bends and dihedrals --------------------------------------------------------
# One constraint alone, unequal inverse masses, compliance > 0, 2 substeps x 3 iterations: a float64
# XPBD run with the finite-difference gradient of the angle -- delta lambda = -(C + alpha lambda) /
# (sum w |grad C|^2 + alpha), x_k += relaxation w_k delta lambda grad_k C, lambda zeroed each substep
# (physics_soft_solve_bend_constraints.cs.hlsl:20-86, physics_cloth_solve_dihedral_constraints.cs.hlsl
# :20-95, physics_clear_soft_constraint_state.cs.hlsl:14-54) -- pins the gradient scale, alpha, the
# lambda accumulation and reset, the inverse-mass weights and the relaxation factor.

K_SOFT_RELAX = 4 * 6 + 0      # kSoftInternalRelaxation: g_SoftSolverConfig0.x (physics_solver_config.hlsli)


def _bend_theta(p):
    u, v = p[0] - p[1], p[2] - p[1]
    return math.acos(float(np.dot(u, v) / np.linalg.norm(u) / np.linalg.norm(v)))


def _dihedral_theta(p, e0, e1, o0, o1):
    a, b, c, d = p[o0], p[o1], p[e0], p[e1]       # p0 = opposite0, p1 = opposite1, p2 = edge0, p3 = edge1
    edge = d - c
    n0 = np.cross(c - a, d - a)
    n1 = np.cross(d - b, c - b)
    n0 /= np.linalg.norm(n0)
    n1 /= np.linalg.norm(n1)
    return math.atan2(float(np.dot(edge / np.linalg.norm(edge), np.cross(n0, n1))),
                      min(max(float(np.dot(n0, n1)), -1.0), 1.0))


def _fd_grad(f, p, eps=1e-7):
    g = np.zeros_like(p)
    for k in range(p.shape[0]):
        for a in range(3):
            d = np.zeros_like(p)
            d[k, a] = eps
            g[k, a] = (f(p + d) - f(p - d)) / (2 * eps)
    return g


def _xpbd_expected(p0, w, theta, rest, compliance, dt_frame, substeps, iters, relax, wrap=False):
    x = p0.astype(np.float64).copy()
    v = np.zeros_like(x)
    dt = dt_frame / substeps
    alpha = compliance / max(dt * dt, 1e-6)
    for _ in range(substeps):
        prev = x.copy()
        x = x + v * dt * (w[:, None] > 1e-6)
        lam = 0.0
        for _ in range(iters):
            g = _fd_grad(theta, x)
            c = theta(x) - rest
            if wrap:
                c = c - 2 * math.pi * math.floor((c + math.pi) / (2 * math.pi))
            dl = -(c + alpha * lam) / (float((w[:, None] * g * g).sum()) + alpha)
            x = x + relax * w[:, None] * dl * g
            lam += dl
        v = (x - prev) / dt
    return x


def _soft_params(substeps=2, iters=3, relax=0.35):
    p = _params(iters=iters, substeps=substeps)
    p.solver_config[K_SOFT_RELAX] = relax
    return p


def test_bend_xpbd_magnitude(oracle):
    pts = np.array([[-1.0, 0.35, 0.1], [0.0, 0.0, 0.0], [0.9, 0.4, -0.2]], np.float32)
    w = np.array([1.0, 0.5, 2.0])
    b = np.zeros(1, au.BEND_DTYPE)
    b["p0"], b["p1"], b["p2"], b["rest_angle"], b["compliance"] = 0, 1, 2, 2.9, 2e-5
    sc = _solid_scene(pts, 1.0, bends=b)
    sc.pos[:, 3] = w
    osc = oracle.OracleScene(sc)
    p = _soft_params()
    oracle.step(osc, p)
    want = _xpbd_expected(pts, w, _bend_theta, 2.9, 2e-5, p.dt, 2, 3, 0.35)
    moved = want - pts
    assert np.abs(moved).max() > 1e-2
    np.testing.assert_allclose(osc.pos[:, :3], want, rtol=0, atol=2e-6 * np.abs(moved).max() + 1e-6)


@pytest.mark.parametrize("rest, lift", [(0.3, 0.45), (-3.0, -0.9)])
def test_dihedral_xpbd_magnitude(oracle, rest, lift):
    """A folded hinge with unequal masses; the second case's angle minus its rest angle lies
    beyond pi, so the constraint wraps by 2 pi."""
    quad = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [1, 0, 1]], np.float32)
    _, d = au.cloth_constraints(quad, np.array([[0, 1, 3], [0, 3, 2]], np.uint32))
    d["rest_angle"], d["compliance"] = rest, 5e-5
    fold = quad.copy()
    fold[2, 1] = lift
    fold[1, 1] = -0.1
    w = np.array([1.0, 0.25, 2.0, 0.7])
    sc = _solid_scene(fold, 1.0, dihedrals=d)
    sc.pos[:, 3] = w
    osc = oracle.OracleScene(sc)
    e0, e1, o0, o1 = (int(d[k][0]) for k in ("e0", "e1", "o0", "o1"))

    def theta(p):
        return _dihedral_theta(p, e0, e1, o0, o1)
    c0 = theta(fold.astype(np.float64)) - rest
    assert (abs(c0) > math.pi) == (rest < -2)
    p = _soft_params()
    oracle.step(osc, p)
    want = _xpbd_expected(fold, w, theta, rest, 5e-5, p.dt, 2, 3, 0.35, wrap=True)
    moved = want - fold
    assert np.abs(moved).max() > 1e-2
    np.testing.assert_allclose(osc.pos[:, :3], want, rtol=0, atol=2e-6 * np.abs(moved).max() + 1e-6)
