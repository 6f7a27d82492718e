"""Long runs: the 64k dam break for 120 steps (config [0]) on the device and in the oracle, and the
tet-volume drift of a soft block over 120 steps.  The device must match the oracle bit for bit at
the end; the density-constraint residual of the last density pass and the volume drift are
reported (printed)."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

pytestmark = pytest.mark.gpu


def test_dam_break_120_steps_bitexact_and_residual(solver, oracle):
    sc = scenes.dam_break_64k()
    p = cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    osc = oracle.OracleScene(sc)
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    for s in range(120):
        solver.step(p)
        st = oracle.step(osc, p)
        if (s + 1) % 20 == 0:
            pos, _, vel = solver.readback()
            same = bool((pos == osc.pos).all() and (vel == osc.vel).all())
            print(f"step {s + 1}: bitwise {same}, max |C| {st.max_abs_constraint_last:.4g}, "
                  f"mean C {st.mean_constraint_last:.4g}, max neighbors {st.max_required_neighbors}", flush=True)
            assert same, s
    m = solver.meta(check=True)
    assert m.fluid_neighbor_truncations == 0
    pos, _, _ = solver.readback()
    r = sc.radii[0]
    assert (pos[:, 1] >= r - 1e-6).all() and (pos[:, 0] <= 4 - r + 1e-6).all()


def _volume(pos, tets):
    p = pos[:, :3].astype(np.float64)
    a, b, c, d = (p[tets["p"][:, k]] for k in range(4))
    return float(np.abs(np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a)).sum() / 6.0)


def test_tet_volume_drift_120_steps(solver):
    sc = scenes.compose([scenes.soft_block_part(8, 0.05, compliance=1e-6)], cnhip.fluid_material(0.03),
                        scenes.container((0, 0, 0), (4, 2, 2)), "block")
    p = cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    v0 = float(sc.tets["rest_volume"].astype(np.float64).sum())
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    for _ in range(120):
        solver.step(p)
    pos, _, _ = solver.readback()
    drift = (_volume(pos, sc.tets) - v0) / v0
    print(f"tet volume drift over 120 steps: {drift:.3e}", flush=True)
    assert abs(drift) < 0.02


def _contacts_between_parts(pos, radii, part):
    """Pairs of particles of different parts closer than their radius sum."""
    from scipy.spatial import cKDTree
    pairs = cKDTree(pos[:, :3].astype(np.float64)).query_pairs(r=2.0 * float(radii.max()), output_type="ndarray")
    a, b = pairs[:, 0], pairs[:, 1]
    d = np.linalg.norm(pos[a, :3].astype(np.float64) - pos[b, :3], axis=1)
    hit = (part[a] != part[b]) & (d < radii[a].astype(np.float64) + radii[b])
    return int(hit.sum())


def test_mixed_full_size_120_steps(solver):
    """The mixed scene at full size (64^3 fluid block, 32^3 tet block of 148,955 tets, 256 x 256
    cloth) for 120 steps: every particle finite and inside the container, the tet block resting on
    the floor with its volume kept, and the cloth landed on the fluid (contacts between parts)."""
    sc = scenes.mixed()
    p = cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)
    info = sc.info
    part = np.repeat(np.arange(3), [info["soft"], info["cloth"], info["fluid"]])
    v0 = float(sc.tets["rest_volume"].astype(np.float64).sum())
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    counts = []
    for s in range(120):
        solver.step(p)
        if (s + 1) % 30 == 0:
            pos, _, vel = solver.readback()
            assert np.isfinite(pos).all() and np.isfinite(vel).all(), s
            c = _contacts_between_parts(pos, sc.radii, part)
            counts.append(c)
            m = solver.meta()
            print(f"mixed step {s + 1}: contacts between parts {c}, max fluid neighbors "
                  f"{m.max_fluid_neighbor_count}, truncations {m.fluid_neighbor_truncations}", flush=True)
    pos, _, _ = solver.readback()
    lo, hi = pos[:, :3].min(0), pos[:, :3].max(0)
    assert (lo >= -1e-3).all() and (hi <= np.array([4.0, 2.0, 2.0]) + 1e-3).all(), (lo, hi)
    soft = pos[part == 0]
    r = float(sc.radii[0])
    drift = (_volume(pos, sc.tets) - v0) / v0
    print(f"mixed: tet volume drift over 120 steps {drift:.3e}, tet block lowest y {soft[:, 1].min():.5f} "
          f"(radius {r})", flush=True)
    assert abs(drift) < 0.02
    assert r - 2e-3 < soft[:, 1].min() < r + 1e-3
    assert counts[-1] > 0
