"""Pins of the oracle's vorticity-confinement passes (CPU only).

physics_fluid_compute_vorticity.cs.hlsl:16-81 and physics_fluid_apply_vorticity_confinement
.cs.hlsl:16-94 have no golden vectors in the reference's tests (their GPU tests only switch the
material term on), so both are pinned by closed forms:

- a rigidly rotating velocity field v = W x x gives, for ANY neighbor set,
  (v_j - v_i) x gradW(d) = (dW/r) (W r^2 - d (d . W)) with d = x_i - x_j, so the curl of every
  particle, interior or at the edge of the lattice, is that sum; a swapped cross product, a lost
  v_i term (edge particles have sum gradW != 0) or a wrong gradient sign fails it;
- on a full cubic lattice, neighbor curl magnitudes a + b . x_j give
  eta = -(1/3) sum(dW r) b, so the confinement direction is exactly b / |b| and
  v_i += dt c (b/|b| x curl_i); a transposed cross product, a missing normalisation or a wrong
  sign fails it.
"""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

F = np.float32


def _lattice(side=9, s=0.01, origin=1.0):
    g = np.arange(side, dtype=np.float32) * F(s) + F(origin)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float32)


def _scene(pts, h, conf, mats=None):
    sc = scenes._fluid_scene(pts, 0.003, np.zeros(0, cnhip.COLLIDER_DTYPE), "vort", {})
    fm = sc.fluid_mats
    if mats is not None:
        fm = np.concatenate([fm] * (int(mats.max()) + 1))
        sc.fluid_mats = fm
        sc.fluid_mat = mats.astype(np.uint32)
    sc.fluid_mats["smoothing_radius"] = h
    sc.fluid_mats["vorticity_confinement"] = conf
    return sc


def _lists(pts, h, inv_mass=None):
    # the list builder's filter: in radius, not itself, not pinned (physics_fluid_build_neighbors
    # .cs.hlsl:88-131)
    d = np.linalg.norm(pts[:, None, :].astype(np.float64) - pts[None, :, :], axis=2)
    live = np.ones(len(pts), bool) if inv_mass is None else inv_mass > 1e-6
    nb = [np.nonzero((d[i] < h) & (np.arange(len(pts)) != i) & live)[0] for i in range(len(pts))]
    cap = max(len(x) for x in nb)
    lists = np.zeros((len(pts), cap), np.uint32)
    counts = np.zeros(len(pts), np.uint32)
    for i, x in enumerate(nb):
        lists[i, :len(x)] = x
        counts[i] = len(x)
    return counts, lists


def _dw(r, h):
    return -2.0 * (15.0 / (np.pi * h ** 3)) * (1.0 - r / h) / h


def test_curl_of_rigid_rotation_closed_form(oracle):
    h = 0.035
    pts = _lattice()
    n = len(pts)
    mats = (np.arange(n) % 3 == 0).astype(np.uint32)  # two materials interleaved
    sc = _scene(pts, h, 0.5, mats)
    sc.fluid_mats["vorticity_confinement"][1] = 0.0      # material 1 does not confine
    om = np.array([0.3, -0.7, 1.1])
    sc.vel[:, :3] = np.cross(om, pts.astype(np.float64)).astype(np.float32)
    sc.pos[5, 3] = 0.0                                   # a pinned particle stores zero curl
    osc = oracle.OracleScene(sc)
    counts, lists = _lists(pts, h, osc.pos[:, 3])
    vort = oracle.fluid_vorticity(osc, counts, lists)
    p64 = osc.pos[:, :3].astype(np.float64)
    checked = 0
    for i in range(n):
        if mats[i] == 1 or i == 5:
            assert np.all(vort[i] == 0.0)
            continue
        want = np.zeros(3)
        for j in lists[i, :counts[i]]:
            if mats[j] != mats[i]:
                continue
            d = p64[i] - p64[j]
            r = np.linalg.norm(d)
            want += _dw(r, h) / r * (om * r * r - d * d.dot(om))
        assert vort[i, :3] == pytest.approx(want, rel=2e-3, abs=1e-3 * np.abs(want).max())
        assert vort[i, 3] == pytest.approx(np.linalg.norm(vort[i, :3].astype(np.float64)), rel=1e-6)
        checked += 1
    assert checked > 400
    # interior particles: the cubic symmetry reduces the sum to (2/3) sum(dW r) W, antiparallel to W
    centre = int(np.argmin(np.linalg.norm(pts - pts.mean(0), axis=1)))
    assert mats[centre] == 0
    assert np.dot(vort[centre, :3], om) < 0


def test_confinement_direction_closed_form(oracle):
    h, conf, dt = 0.035, 0.7, F(1.0 / 240.0)
    pts = _lattice()
    n = len(pts)
    sc = _scene(pts, h, conf)
    sc.vel[:, :3] = F(0.25)
    sc.vel[:, 3] = F(9.0)
    osc = oracle.OracleScene(sc)
    counts, lists = _lists(pts, h)
    a, b = 3.0, np.array([40.0, -15.0, 25.0])
    vort = np.zeros((n, 4), np.float32)
    vort[:, 3] = (a + (pts.astype(np.float64) - 1.0) @ b).astype(np.float32)
    centre = int(np.argmin(np.linalg.norm(pts - pts.mean(0), axis=1)))
    curl = np.array([0.2, 1.5, -0.6], np.float32)
    vort[centre, :3] = curl
    vort[centre, 3] = np.linalg.norm(curl)
    quiet = 0                                   # a corner outside the centre's radius, |curl| = 0
    assert np.linalg.norm(pts[quiet] - pts[centre]) > h
    vort[quiet, :3] = 0.0
    vort[quiet, 3] = 0.0
    v0 = osc.vel.copy()
    oracle.fluid_vorticity_apply(osc, dt, counts, lists, vort)
    bn = b / np.linalg.norm(b)
    want = v0[centre, :3].astype(np.float64) + float(dt) * conf * np.cross(bn, curl.astype(np.float64))
    assert osc.vel[centre, :3] == pytest.approx(want, rel=1e-5, abs=1e-6)
    assert osc.vel[centre, 3] == 0.0
    assert np.array_equal(osc.vel[quiet], v0[quiet])
    assert not np.array_equal(osc.vel[centre], v0[centre])


def test_step_applies_confinement_only_when_enabled(oracle):
    # a rotating blob: the substep with confinement differs from the one without only in velocity
    pts = _lattice(side=8, s=0.02)
    sc = _scene(pts, 0.05, 0.0)
    sc.vel[:, :3] = np.cross([0.0, 0.0, 2.0], pts.astype(np.float64) - pts.mean(0)).astype(np.float32)
    p = cnhip.default_params(dt=1 / 120, substeps=1, default_iterations=2, max_fluid_neighbors=96,
                             gravity=(0.0, 0.0, 0.0, 0.0))
    off = oracle.OracleScene(sc)
    oracle.step(off, p)
    sc.fluid_mats["vorticity_confinement"] = 0.4
    on = oracle.OracleScene(sc)
    oracle.step(on, p)
    assert np.array_equal(on.pos, off.pos)
    assert np.array_equal(on.prev, off.prev)
    assert not np.array_equal(on.vel, off.vel)
    assert np.all(on.vel[:, 3] == 0.0)


# ---- render anisotropy (physics_fluid_build_render_anisotropy.cs.hlsl:34-171) ----

def _aniso_scene(pts, radius, h, kinds=None):
    sc = scenes._fluid_scene(pts, radius, np.zeros(0, cnhip.COLLIDER_DTYPE), "aniso", {})
    sc.fluid_mats["smoothing_radius"] = h
    if kinds is not None:
        sc.kinds = kinds.astype(np.uint32)
    return sc


def test_anisotropy_isotropic_cases(oracle):
    # a lone fluid particle, a solid particle and a pinned fluid particle: identity frame, scale r
    pts = np.array([[0.0, 0.0, 0.0], [5.0, 0.0, 0.0], [0.01, 0.0, 0.0], [9.0, 9.0, 9.0]], np.float32)
    kinds = np.array([1, 1, 0, 1])
    sc = _aniso_scene(pts, 0.004, 0.05, kinds)
    sc.pos[3, 3] = 0.0
    osc = oracle.OracleScene(sc)
    counts, lists = _lists(pts, 0.05, osc.pos[:, 3])
    counts[3] = 7  # a pinned particle's stale count is never read
    out = oracle.fluid_render_anisotropy(osc, counts, lists)
    for i in (1, 2, 3):
        np.testing.assert_array_equal(out[:, i, :3], np.eye(3, dtype=np.float32))
        assert np.all(out[:, i, 3] == F(0.004))


def test_anisotropy_of_a_line_and_a_plane(oracle):
    r, h, s = 0.01, 0.1, 0.012
    u = np.array([1.0, 2.0, 2.0]) / 3.0
    t = np.array([-4, -3, -2, -1, 1, 2, 3, 4], np.float64) * s
    line = np.concatenate([[[0.5, 0.5, 0.5]], 0.5 + t[:, None] * u[None, :]]).astype(np.float32)
    g = np.arange(-3, 4) * s
    gx, gy = np.meshgrid(g, g)
    plane = np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], 1)
    plane = plane[np.argsort(np.abs(plane).sum(1), kind="stable")] + np.array([2.0, 2.0, 2.0])
    pts = np.concatenate([line, plane]).astype(np.float32)
    sc = _aniso_scene(pts, r, h)
    osc = oracle.OracleScene(sc)
    counts, lists = _lists(pts, h)
    out = oracle.fluid_render_anisotropy(osc, counts, lists)
    # the line: one axis along u with 1.1 sqrt(weighted variance); the other scales clamp to r
    p = osc.pos[:9, :3].astype(np.float64)
    d = np.linalg.norm(p[1:] - p[0], axis=1)
    w = 1.0 - (d / h) ** 3
    var = (w * d ** 2).sum() / w.sum()
    assert abs(abs(out[0, 0, :3].astype(np.float64) @ u) - 1.0) < 1e-5
    assert out[0, 0, 3] == pytest.approx(min(max(1.1 * np.sqrt(var), r), 2 * r), rel=1e-5)
    assert out[1, 0, 3] == F(r) and out[2, 0, 3] == F(r)
    # the plane z = 2 (centre is its first point): the last axis is the normal, scales descend
    c = 9
    assert abs(abs(out[2, c, 2]) - 1.0) < 1e-6
    assert out[0, c, 3] >= out[1, c, 3] >= out[2, c, 3] == F(r)
    q = osc.pos[9:, :3].astype(np.float64)
    dd = np.linalg.norm(q - q[0], axis=1)
    m = (dd > 0) & (dd < h)
    ww = 1.0 - (dd[m] / h) ** 3
    mean = (ww[:, None] * q[m]).sum(0) / ww.sum()
    cov = ((ww[:, None, None] * (q[m] - mean)[:, :, None] * (q[m] - mean)[:, None, :]).sum(0)) / ww.sum()
    lam = np.sort(np.linalg.eigvalsh(cov))[::-1]
    want = np.clip(1.1 * np.sqrt(np.maximum(lam, 0)), r, 2 * r)
    assert out[:, c, 3] == pytest.approx(want, rel=1e-4)
    # every frame is orthonormal
    for i in (0, c):
        ax = out[:, i, :3].astype(np.float64)
        np.testing.assert_allclose(ax @ ax.T, np.eye(3), atol=1e-5)


def test_step_writes_the_anisotropy_of_the_final_positions(oracle):
    pts = _lattice(side=6, s=0.02)
    sc = _scene(pts, 0.05, 0.0)
    sc.radii[:] = F(0.02)  # scales clamp to [0.02, 0.04]: the lattice's spread lands inside
    p = cnhip.default_params(dt=1 / 120, substeps=2, default_iterations=2, max_fluid_neighbors=96)
    osc = oracle.OracleScene(sc)
    out = osc.enable_render_anisotropy()
    oracle.step(osc, p)
    counts, lists = _lists(osc.pos[:, :3], 0.05)
    again = oracle.fluid_render_anisotropy(osc, counts, lists)
    # the lists here are in index order, the step's in walk order: the scales agree to rounding
    # (axes of near-equal eigenvalues are not determined by the data, so only the scales compare)
    np.testing.assert_allclose(out[:, :, 3], again[:, :, 3], rtol=1e-4)
    assert np.any(out[:, :, 3] != out[0, 0, 3])
