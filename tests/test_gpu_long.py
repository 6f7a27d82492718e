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
