"""Pins of the oracle's fluid position-correction and velocity passes (CPU only).

Each case builds a few particles by hand, runs one oracle pass and compares it with a float64
evaluation of the reference shader's formula written out here term by term:

- the delta pass, physics_fluid_compute_delta_positions.cs.hlsl:27-157 (pressure
  -1/2 (C_i + C_j) W'(r) n_ij, cohesion, viscosity, the CFL clamp on both sides of -cfl, surface
  tension, same-material gating, skipped neighbors) and its static-wall term
  (ComputeFluidBoundaryDeltaFromCandidates, physics_fluid_boundary_common.hlsli:140-167);
- the relaxation division Delta / max(w relax, 1), physics_fluid_apply_delta_positions.cs.hlsl:11-31;
- the velocity update, physics_particle_update_velocities.cs.hlsl:13-44, and the wall velocity
  projection with friction and tangential damping,
  physics_fluid_project_boundary_velocities.cs.hlsl:26-118 with CombineContactMaterial and
  ComputePositionFrictionDelta (physics_rigid_solver_shared.hlsli:37-67);
- the wall push-out, physics_fluid_clamp_boundary.cs.hlsl:23-80.

Solver-config entries are addressed by the reference cbuffer order (physics_solver_config.hlsli:5-51:
ten float4 slots, contact 0-1, rigid 2-3, joint 4-5, soft 6-7, fluid 8-9) and set to values
unlike their defaults, so a swapped entry changes the result.  A dropped term, a wrong sign, a
missing 1/2 or a transposed index moves the result by far more than the float32 rounding the
tolerances allow.
"""
import math

import numpy as np
import pytest

import cnhip
from cnhip import scenes

F = np.float32
EPS = 1.0e-6                 # kEpsilon, physics_base.hlsli:7
PI_LIT = 3.14159265359       # the literal of physics_fluid_common.hlsli:26

# float index = 4 * slot + component, slots in cbuffer order
SLOT = {"contact0": 0, "contact1": 1, "soft0": 6, "fluid0": 8, "fluid1": 9}
K_CONTACT_SLOP = 4 * SLOT["contact0"] + 0
K_POS_FRICTION_MIN = 4 * SLOT["contact1"] + 3
K_FLUID_SOLVE = 4 * SLOT["fluid0"] + 0
K_FLUID_RELAX = 4 * SLOT["fluid0"] + 1
K_FLUID_BND_DENSITY = 4 * SLOT["fluid0"] + 2
K_FLUID_BND_DELTA = 4 * SLOT["fluid0"] + 3

H = 0.1
DT = 1.0 / 480.0


def cfg_values(solve=0.7, relax=0.25, bnd_delta=0.3, slop=0.004, fric_min=1e-5):
    c = np.array(list(cnhip.default_params().solver_config), np.float32)
    c[K_FLUID_SOLVE] = solve
    c[K_FLUID_RELAX] = relax
    c[K_FLUID_BND_DELTA] = bnd_delta
    c[K_FLUID_BND_DENSITY] = 0.5
    c[K_CONTACT_SLOP] = slop
    c[K_POS_FRICTION_MIN] = fric_min
    return c


def material(h=H, viscosity=0.0, cohesion=0.0, c1=0.0, c2=0.0, tension=0.0, cfl=0.0):
    m = np.zeros(1, cnhip.FLUID_MATERIAL_DTYPE)
    m["rest_density"] = 1000.0
    m["inv_rest_density"] = 1e-3
    m["smoothing_radius"] = h
    m["density_constraint_scale"] = 1.0
    m["viscosity"] = viscosity
    m["cohesion"] = cohesion
    m["cohesion1"] = c1
    m["cohesion2"] = c2
    m["surface_tension"] = tension
    m["cfl_radius"] = cfl
    return m


def box(center, half, material=(0.0, 0.0, 0.0, 0.0)):
    c = np.zeros(1, cnhip.COLLIDER_DTYPE)
    c["center"], c["half_extents"], c["material"] = center, half, material
    c["layer"], c["mask"], c["enabled"] = 1, 0xFFFFFFFF, 1
    return c


def fluid(points, mats, mat_idx=None, colliders=None, radius=0.03):
    pts = np.asarray(points, np.float32)
    cols = np.zeros(0, cnhip.COLLIDER_DTYPE) if colliders is None else np.concatenate(colliders)
    sc = scenes._fluid_scene(pts, radius, cols, "pin", {})
    sc.fluid_mats = np.concatenate(mats)
    if mat_idx is not None:
        sc.fluid_mat = np.asarray(mat_idx, np.uint32)
    return sc


# ---- float64 restatements of the shader formulas -------------------------------------------------

def w_spiky(d, h):
    return 0.0 if d >= h else 15.0 / (PI_LIT * h ** 3) * (1.0 - d / h) ** 2


def dw_spiky(d, h):
    if d <= EPS or d >= h:
        return 0.0
    return -2.0 * 15.0 / (PI_LIT * h ** 3) * (1.0 - d / h) / h


def k_cohesion(d, h, c1, c2):
    if d <= EPS or d >= h:
        return 0.0
    q = d / h
    return c1 * q ** 3 + c2 * q ** 2 - 1.0


def delta_expected(pos, dacc, nc, m_of, mats, i, neigh, cfg, dt, walls=(), iteration=1):
    """physics_fluid_compute_delta_positions.cs.hlsl:27-157 in float64 for particle i."""
    f64 = np.float64
    di = dacc[i, :3].astype(f64) if iteration > 0 else np.zeros(3)
    xi = pos[i, :3].astype(f64) + di
    m = mats[m_of[i]]
    h = max(float(m["smoothing_radius"]), 1e-4)
    coh = dt * max(float(m["cohesion"]), 0.0)
    vis = dt * max(float(m["viscosity"]), 0.0)
    st = max(float(m["surface_tension"]), 0.0)
    cfl = max(float(m["cfl_radius"]), 0.0)
    Ci = float(nc[i, 3])
    sn_i = nc[i, :3].astype(f64) if st > EPS else np.zeros(3)
    acc, w = np.zeros(3), 0.0
    for j in neigh:
        if pos[j, 3] <= EPS:
            continue
        dj = dacc[j, :3].astype(f64) if iteration > 0 else np.zeros(3)
        d = xi - (pos[j, :3].astype(f64) + dj)
        r = float(np.linalg.norm(d))
        if r >= h or r <= EPS:
            continue
        n = d / r
        pair = -0.5 * (Ci + float(nc[j, 3])) * dw_spiky(r, h) * n
        if m_of[i] == m_of[j]:
            rel = di - dj
            if coh > EPS:
                pair = pair - n * (coh * k_cohesion(r, h, float(m["cohesion1"]), float(m["cohesion2"])))
            if vis > EPS:
                pair = pair - (1.0 - 1.0 / (1.0 + vis * w_spiky(r, h))) * rel
            if cfl > EPS:
                proj = float(rel @ n)
                if proj < -cfl:
                    pair = pair - n * (proj + cfl) * 0.5
            if st > EPS:
                pair = pair - dt * (sn_i - nc[j, :3].astype(f64))
        acc += pair
        w += 1.0
    bd = np.zeros(3)
    for center, half in walls:   # axis-aligned boxes; the particle outside them
        local = xi - np.asarray(center, f64)
        dv = local - np.clip(local, -np.asarray(half, f64), np.asarray(half, f64))
        sd = float(np.linalg.norm(dv))
        if sd <= EPS or sd >= h:     # inside (or on) the box, or beyond the support: no ghost
            continue
        g = dv / sd * (2.0 * sd)
        gl = float(np.linalg.norm(g))
        if gl <= EPS or gl >= h:
            continue
        bd += -max(Ci, 0.0) * dw_spiky(gl, h) * (g / gl)
    return acc * float(cfg[K_FLUID_SOLVE]) + bd * float(cfg[K_FLUID_BND_DELTA]), w


def close(got, want, rel=4e-6):
    scale = max(float(np.abs(want).max()), 1e-30)
    np.testing.assert_allclose(np.asarray(got, np.float64), want, rtol=0, atol=rel * scale + 1e-12)


# ---- A9: the delta pass ---------------------------------------------------------------------------

def neighborhood(seed=0):
    """Particle 0 and four listed neighbors: two in radius (0.35 h, 0.72 h), one beyond it
    (1.02 h, listed but skipped), one pinned (inverse mass 0, skipped)."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(4, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    x0 = np.array([0.5, 0.5, 0.5])
    pts = [x0] + [x0 + r * H * u[k] for k, r in enumerate((0.35, 0.72, 1.02, 0.5))]
    return np.array(pts)


CASES = {
    "pressure": dict(),
    "cohesion": dict(cohesion=3.0, c1=2.5, c2=-1.3),
    "viscosity": dict(viscosity=40.0),
    "cfl_active": dict(cfl=1e-3),
    "cfl_inactive": dict(cfl=5e-2),
    "tension": dict(tension=0.6),
    "all_terms": dict(cohesion=3.0, c1=2.5, c2=-1.3, viscosity=40.0, cfl=1e-3, tension=0.6),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("iteration", [0, 1])
def test_delta_pass_terms(oracle, case, iteration):
    pts = neighborhood(seed=3)
    mats = [material(**CASES[case])]
    sc = fluid(pts, mats)
    sc.pos[4, 3] = 0.0                                   # pinned neighbor
    osc = oracle.OracleScene(sc)
    n = sc.n
    nc = np.zeros((n, 4), np.float32)
    # C_i, C_j of the size a density pass gives (C W' of about 1e-4 m), so that every term counts
    nc[:, 3] = [3.1e-9, -1.2e-9, 0.8e-9, 5e-9, 9e-9]
    nc[:, :3] = np.array([[0.2, -0.1, 0.05], [-0.3, 0.4, 0.1], [0.0, 0.2, -0.25], [1, 1, 1], [1, 1, 1]])
    dacc = np.zeros((n, 4), np.float32)
    # relative deltas chosen so the CFL projection (rel . n) is about -4e-3 for the first neighbor
    dacc[0, :3] = (pts[0] - pts[1]) / np.linalg.norm(pts[0] - pts[1]) * -0.003
    dacc[1, :3] = (pts[0] - pts[1]) / np.linalg.norm(pts[0] - pts[1]) * 0.001
    dacc[2, :3] = [0.0007, -0.0004, 0.0002]
    counts = np.zeros(n, np.uint32)
    lists = np.zeros((n, 8), np.uint32)
    counts[0] = 4
    lists[0, :4] = [1, 2, 3, 4]
    cfg = cfg_values()
    out = oracle.fluid_delta(osc, cfg, DT, counts, lists, nc, dacc=dacc, iteration=iteration)
    want, w = delta_expected(sc.pos, dacc, nc, sc.fluid_mat, sc.fluid_mats, 0, [1, 2, 3, 4], cfg, DT,
                             iteration=iteration)
    assert out[0, 3] == w == 2.0                          # two neighbors interact
    close(out[0, :3], want)
    assert not out[1:].any()                              # nothing listed for the others
    # each optional term contributes (so the comparison above tests it): against the pressure-only
    # material the result moves by far more than the tolerance, except where the term is off
    sc2 = fluid(pts, [material()])
    sc2.pos[4, 3] = 0.0
    base = oracle.fluid_delta(oracle.OracleScene(sc2), cfg, DT, counts, lists, nc, dacc=dacc, iteration=iteration)
    moved = np.abs(base[0, :3] - out[0, :3]).max() / np.abs(out[0, :3]).max()
    inert = case in ("pressure", "cfl_inactive") or (iteration == 0 and case in ("viscosity", "cfl_active"))
    if inert:
        assert moved == 0.0
    else:
        assert moved > 1e-2, moved


def test_delta_pass_same_material_gating(oracle):
    """A neighbor of another fluid material adds only the pressure term."""
    pts = neighborhood(seed=5)
    kw = CASES["all_terms"]
    sc = fluid(pts[:3], [material(**kw), material(**kw)], mat_idx=[0, 0, 1])
    osc = oracle.OracleScene(sc)
    nc = np.zeros((3, 4), np.float32)
    nc[:, 3] = [2e-9, 1e-9, -0.5e-9]
    nc[:, :3] = [[0.1, 0.2, 0.3], [-0.2, 0.0, 0.1], [0.3, -0.3, 0.0]]
    dacc = np.zeros((3, 4), np.float32)
    dacc[:, :3] = [[0.001, 0.0, -0.002], [-0.001, 0.002, 0.0], [0.0005, 0.0005, 0.0005]]
    counts = np.array([2, 0, 0], np.uint32)
    lists = np.zeros((3, 4), np.uint32)
    lists[0, :2] = [1, 2]
    cfg = cfg_values()
    out = oracle.fluid_delta(osc, cfg, DT, counts, lists, nc, dacc=dacc, iteration=1)
    want, w = delta_expected(sc.pos, dacc, nc, sc.fluid_mat, sc.fluid_mats, 0, [1, 2], cfg, DT)
    close(out[0, :3], want)
    # the cross-material neighbor's share is exactly its pressure term
    only1, _ = delta_expected(sc.pos, dacc, nc, sc.fluid_mat, sc.fluid_mats, 0, [1], cfg, DT)
    f64 = np.float64
    r = (sc.pos[0, :3].astype(f64) + dacc[0, :3]) - (sc.pos[2, :3].astype(f64) + dacc[2, :3])
    d = float(np.linalg.norm(r))
    press = -0.5 * (2e-9 - 0.5e-9) * dw_spiky(d, H) * r / d * float(cfg[K_FLUID_SOLVE])
    close(want - only1, press, rel=1e-9)


def test_delta_pass_pair_antisymmetry(oracle):
    """Two particles with only the pressure term push each other apart by equal and opposite
    amounts, 1/2 (C_i + C_j) |W'(d)| each, along the line between them."""
    d = 0.4 * H
    sc = fluid([[0.5, 0.5, 0.5], [0.5 + d, 0.5, 0.5]], [material()])
    osc = oracle.OracleScene(sc)
    nc = np.zeros((2, 4), np.float32)
    nc[:, 3] = [2.5e-9, 1.5e-9]
    counts = np.array([1, 1], np.uint32)
    lists = np.array([[1], [0]], np.uint32)
    cfg = cfg_values(solve=1.0)
    out = oracle.fluid_delta(osc, cfg, DT, counts, lists, nc)
    mag = 0.5 * 4e-9 * abs(dw_spiky(float(np.float32(0.5 + d) - np.float32(0.5)), H))
    assert out[0, 0] < 0 < out[1, 0]                      # apart
    assert abs(out[0, 0]) == pytest.approx(mag, rel=2e-6)
    assert out[0, 0] == -out[1, 0]
    assert not out[:, 1:3].any()


@pytest.mark.parametrize("ci", [4e-9, -4e-9])
def test_delta_pass_single_wall(oracle, ci):
    """A particle 0.3 h above a floor: ghost 2 sd n, term -max(C_i, 0) W'(2 sd) n times the
    boundary delta scale (zero for a negative C_i); a particle inside the box or beyond h gets
    no ghost."""
    floor_c, floor_h = (0.5, -0.5, 0.5), (1.0, 0.5, 1.0)      # top face at y = 0
    pts = [[0.5, 0.3 * H, 0.5], [0.2, -0.01, 0.2], [0.8, 1.2 * H, 0.8]]
    sc = fluid(pts, [material()], colliders=[box(floor_c, floor_h)])
    osc = oracle.OracleScene(sc)
    nc = np.zeros((3, 4), np.float32)
    nc[:, 3] = ci
    counts = np.zeros(3, np.uint32)
    lists = np.zeros((3, 2), np.uint32)
    br = np.array([[0, 1], [0, 1], [0, 1]], np.uint32)
    bp = np.array([0], np.uint32)
    cfg = cfg_values()
    out = oracle.fluid_delta(osc, cfg, DT, counts, lists, nc, bnd_ranges=br, bnd_pairs=bp)
    for i in range(3):
        want, _ = delta_expected(sc.pos, np.zeros((3, 4), np.float32), nc, sc.fluid_mat, sc.fluid_mats, i, [], cfg,
                                 DT, walls=[(floor_c, floor_h)], iteration=0)
        close(out[i, :3], want)
    if ci > 0:
        sd = float(np.float32(0.3 * H))
        assert out[0, 1] == pytest.approx(-ci * dw_spiky(2 * sd, H) * 0.3, rel=4e-6)
        assert out[0, 0] == 0.0 and out[0, 2] == 0.0
    assert not out[1:].any()
    assert not out[:, 3].any()                            # walls add no interaction weight


def test_delta_pass_corner_sums_two_walls(oracle):
    fl = ((0.5, -0.5, 0.5), (1.0, 0.5, 1.0))                   # floor, top at y = 0
    wall = ((-0.5, 0.5, 0.5), (0.5, 1.0, 1.0))                 # x-low wall, face at x = 0
    pts = [[0.2 * H, 0.35 * H, 0.5]]
    sc = fluid(pts, [material()], colliders=[box(*fl), box(*wall)])
    osc = oracle.OracleScene(sc)
    nc = np.array([[0, 0, 0, 3e-9]], np.float32)
    br = np.array([[0, 2]], np.uint32)
    bp = np.array([0, 1], np.uint32)
    cfg = cfg_values()
    out = oracle.fluid_delta(osc, cfg, DT, np.zeros(1, np.uint32), np.zeros((1, 1), np.uint32), nc,
                             bnd_ranges=br, bnd_pairs=bp)
    want, _ = delta_expected(sc.pos, np.zeros((1, 4), np.float32), nc, sc.fluid_mat, sc.fluid_mats, 0, [], cfg, DT,
                             walls=[fl, wall], iteration=0)
    close(out[0, :3], want)
    assert out[0, 0] > 0 and out[0, 1] > 0                 # pushed away from both walls


# ---- A10: the relaxation division ----------------------------------------------------------------

@pytest.mark.parametrize("iteration", [0, 1])
def test_apply_relaxation(oracle, iteration):
    """Delta / max(w relax, 1): weights below 1/relax divide by 1, above it by w relax; iteration 0
    starts the accumulation from zero, later iterations add; solids are untouched."""
    relax = 0.25
    sc = fluid([[0.1, 0.1, 0.1], [0.3, 0.3, 0.3], [0.5, 0.5, 0.5], [0.7, 0.7, 0.7]], [material()])
    sc.kinds[3] = cnhip.CN_PARTICLE_KIND_SOLID
    osc = oracle.OracleScene(sc)
    din = np.array([[0.004, -0.002, 0.001, 2.0],       # w relax = 0.5 -> / 1
                    [0.004, -0.002, 0.001, 12.0],      # w relax = 3   -> / 3
                    [0.004, -0.002, 0.001, 4.0],       # w relax = 1   -> / 1
                    [0.004, -0.002, 0.001, 12.0]], np.float32)
    prev = np.array([[0.01, 0.02, 0.03, 7.0]] * 4, np.float32)
    acc = oracle.fluid_apply(osc, cfg_values(relax=relax), iteration, din, prev)
    base = prev[:, :3].astype(np.float64) if iteration > 0 else np.zeros((4, 3))
    for i, div in enumerate((1.0, 3.0, 1.0)):
        np.testing.assert_allclose(acc[i, :3], base[i] + din[i, :3].astype(np.float64) / div, rtol=1e-6, atol=1e-9)
        assert acc[i, 3] == 0.0
    np.testing.assert_array_equal(acc[3], prev[3])         # the solid keeps its buffer


def test_clamp_pushes_to_radius_in_a_corner(oracle):
    """Two walls, a particle penetrating both: the sweeps push it out along each normal to exactly
    its radius from each face (physics_fluid_clamp_boundary.cs.hlsl:23-80)."""
    r = 0.03
    fl = box((0.5, -0.5, 0.5), (1.0, 0.5, 1.0))
    wall = box((-0.5, 0.5, 0.5), (0.5, 1.0, 1.0))
    sc = fluid([[0.01, 0.005, 0.5], [0.5, 0.5, 0.5]], [material()], colliders=[fl, wall], radius=r)
    osc = oracle.OracleScene(sc)
    dacc = np.zeros((2, 4), np.float32)
    dacc[0, :3] = [-0.002, 0.001, 0.0]
    br = np.array([[0, 2], [0, 2]], np.uint32)
    out = oracle.fluid_clamp(osc, dacc, br, np.array([0, 1], np.uint32))
    cur = sc.pos[0, :3] + out[0, :3]
    assert cur[0] == pytest.approx(r, abs=1e-7) and cur[1] == pytest.approx(r, abs=1e-7)
    assert cur[2] == sc.pos[0, 2]
    np.testing.assert_array_equal(out[1, :3], dacc[1, :3])     # far from both: unchanged


# ---- A13: velocity update and wall velocity projection -------------------------------------------

def test_velocity_update_with_damping(oracle):
    """v = (x* + Delta - x_prev) / dt * max(1 - damping dt, 0); fluid positions take the delta and
    the delta buffer is cleared; a pinned particle's velocity is zero."""
    dt = 1 / 240
    sc = fluid([[0.1, 0.2, 0.3], [0.4, 0.5, 0.6], [0.7, 0.8, 0.9]], [material()])
    sc.part_mats = np.array([[0.0, 0.0, 30.0, 0.0]], np.float32)    # damping 30: scale 0.875
    sc.pos[2, 3] = 0.0
    osc = oracle.OracleScene(sc)
    osc.prev[:, :3] = osc.pos[:, :3] - np.array([[0.001, -0.002, 0.0005]], np.float32)
    dacc = np.zeros((3, 4), np.float32)
    dacc[:, :3] = [[0.0002, 0.0001, -0.0003], [0.0, 0.0, 0.0], [0.1, 0.1, 0.1]]
    x0 = osc.pos[:, :3].astype(np.float64).copy()
    left = oracle.update_velocities(osc, dt, dacc)
    assert not left[:2].any()
    for i in range(2):
        x = x0[i] + dacc[i, :3]
        np.testing.assert_allclose(osc.pos[i, :3], x, rtol=0, atol=1e-7)
        want = (x - osc.prev[i, :3].astype(np.float64)) / dt * max(1 - 30.0 * dt, 0.0)
        np.testing.assert_allclose(osc.vel[i, :3], want, rtol=2e-4, atol=1e-6)
    assert not osc.vel[2].any()


def combine(pm, cm):
    """CombineContactMaterial: (kinetic, restitution, static)."""
    return (math.sqrt(max(0, pm[0]) * max(0, cm[0])), max(max(0, pm[1]), max(0, cm[1])),
            math.sqrt(max(0, pm[3]) * max(0, cm[3])))


def project_expected(v, sd, normal, r, pm, cm, cfg, dt):
    """physics_fluid_project_boundary_velocities.cs.hlsl:26-118 for one wall contact, float64."""
    slop = float(cfg[K_CONTACT_SLOP])
    n = np.asarray(normal, np.float64)
    kin, _, stat = combine(pm, cm)
    pen = max(r - sd, 0.0)
    cd = max(r + 2 * slop, 1e-5)
    cw = min(max((cd - sd) / cd, 0.0), 1.0)
    rv = np.asarray(v, np.float64).copy()
    nv = rv @ n
    if nv < 0:
        rv -= n * nv
    tv = rv - n * (rv @ n)
    if dt > EPS:
        disp = tv * dt
        td = float(np.linalg.norm(disp))
        p = max(pen, slop)
        fd = np.zeros(3)
        if not (td <= float(cfg[K_POS_FRICTION_MIN]) or p <= EPS):
            if td <= min(max(stat, 0), 1) * p:
                fd = disp
            else:
                fd = disp * min(min(max(kin, 0), 1) * p / td, 1.0)
        tv = tv - fd / dt
    tv = tv * max(1.0 - max(pm[2], 0.0) * cw * dt, 0.0)
    return n * max(rv @ n, 0.0) + tv


WALL_CASES = {
    # name: (velocity, height of the particle above the floor, particle material, wall material)
    "static_friction_holds": ((0.3, -1.0, 0.1), 0.02, (0.5, 0.0, 2.0, 0.4), (0.8, 0.0, 0.0, 0.9)),
    "kinetic_friction_slides": ((10.0, -1.0, 3.0), 0.02, (0.5, 0.0, 2.0, 0.4), (0.8, 0.0, 0.0, 0.9)),
    "outward_normal_kept": ((6.0, 0.5, 0.0), 0.025, (0.3, 0.0, 4.0, 0.0), (0.6, 0.0, 0.0, 0.0)),
    "damping_only": ((2.0, -0.2, -1.0), 0.031, (0.0, 0.0, 50.0, 0.0), (0.0, 0.0, 0.0, 0.0)),
    "penetrating": ((5.0, -3.0, 1.0), 0.01, (0.2, 0.0, 1.0, 0.1), (0.9, 0.0, 0.0, 0.2)),
    # within the contact distance r + 2 slop but not touching: friction acts over the slop
    "within_slop": ((3.0, -0.5, -2.0), 0.033, (0.7, 0.0, 1.0, 0.3), (0.7, 0.0, 0.0, 0.3)),
}


@pytest.mark.parametrize("case", list(WALL_CASES))
def test_wall_velocity_projection(oracle, case):
    v, y, pm, cm = WALL_CASES[case]
    r, dt = 0.03, 1 / 480
    cfg = cfg_values(slop=0.004, fric_min=1e-5)
    sc = fluid([[0.5, y, 0.5], [0.5, 0.5, 0.5]], [material()], colliders=[box((0.5, -0.5, 0.5), (1.0, 0.5, 1.0), cm)],
               radius=r)
    sc.part_mats = np.array([pm], np.float32)
    osc = oracle.OracleScene(sc)
    osc.vel[:, :3] = v
    br = np.array([[0, 1], [0, 1]], np.uint32)
    oracle.project_boundary_velocities(osc, cfg, dt, br, np.array([0], np.uint32))
    sd = float(np.float32(y))
    want = project_expected(v, sd, (0, 1, 0), r, pm, cm, cfg, dt)
    np.testing.assert_allclose(osc.vel[0, :3], want, rtol=2e-6, atol=2e-6 * float(np.abs(want).max() + 1e-3))
    assert osc.vel[0, 3] == 0.0
    np.testing.assert_array_equal(osc.vel[1, :3], np.float32(v))   # beyond the contact distance
    if case == "static_friction_holds":
        assert not osc.vel[0, [0, 2]].any() and osc.vel[0, 1] == 0.0
    if case == "kinetic_friction_slides":
        assert 0 < osc.vel[0, 0] < v[0]


def test_wall_projection_takes_the_closest_wall(oracle):
    """Of two walls in contact range, the nearer one's normal and combined material are used."""
    r, dt = 0.03, 1 / 480
    cfg = cfg_values()
    fl = box((0.5, -0.5, 0.5), (1.0, 0.5, 1.0), (0.9, 0.0, 0.0, 0.9))     # floor, y = 0
    wall = box((-0.5, 0.5, 0.5), (0.5, 1.0, 1.0), (0.1, 0.0, 0.0, 0.0))   # x-low wall, x = 0
    pm = (0.5, 0.0, 3.0, 0.5)
    v = (-4.0, -2.0, 1.5)
    for pos, normal, cm in (([0.015, 0.025, 0.5], (1, 0, 0), (0.1, 0.0, 0.0, 0.0)),
                            ([0.028, 0.012, 0.5], (0, 1, 0), (0.9, 0.0, 0.0, 0.9))):
        sc = fluid([pos], [material()], colliders=[fl, wall], radius=r)
        sc.part_mats = np.array([pm], np.float32)
        osc = oracle.OracleScene(sc)
        osc.vel[0, :3] = v
        oracle.project_boundary_velocities(osc, cfg, dt, np.array([[0, 2]], np.uint32), np.array([0, 1], np.uint32))
        sd = float(min(np.float32(pos[0]), np.float32(pos[1])))
        want = project_expected(v, sd, normal, r, pm, cm, cfg, dt)
        np.testing.assert_allclose(osc.vel[0, :3], want, rtol=2e-6, atol=1e-5)
