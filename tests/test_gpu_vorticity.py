"""GPU parity of vorticity confinement (physics_fluid_compute_vorticity.cs.hlsl and
physics_fluid_apply_vorticity_confinement.cs.hlsl after the velocity update) against the oracle:
whole substeps, positions and velocities bit for bit, for a single-material dam break (the
uniform-class walk), a mixed scene where only one of two materials confines (the same-material
test, other environments, solid and pinned particles, masked layers) and a neighbor cap small
enough to truncate the post-update list."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

from test_gpu_fluid_parity import mixed_scene, params

pytestmark = pytest.mark.gpu
F = np.float32


def swirl(sc, rate=3.0):
    c = sc.pos[:, :3].astype(np.float64).mean(0)
    sc.vel[:, :3] = np.cross([0.0, rate, rate], sc.pos[:, :3].astype(np.float64) - c).astype(np.float32)
    return sc


def step_pair(sc, p, steps, solver, oracle):
    osc = oracle.OracleScene(sc)
    solver.reserve(sc.n, max(p.max_fluid_neighbors, 64))
    solver.upload(sc)
    out = []
    for _ in range(steps):
        solver.step(p)
        oracle.step(osc, p)
        pos, _, vel = solver.readback()
        assert np.isfinite(pos).all() and np.isfinite(vel).all()
        out.append(bool((pos == osc.pos).all() and (vel == osc.vel).all()))
    return out, osc


def test_vorticity_dam_break_bitexact(solver, oracle):
    sc = swirl(scenes.dam_break_64k(side=14))
    sc.fluid_mats = cnhip.fluid_material(0.03, vorticity_confinement=0.6)
    res, osc = step_pair(sc, params(cap=64), 5, solver, oracle)
    print(res)
    assert all(res), res
    # the term is live: without it the same steps end elsewhere
    off = sc.copy()
    off.fluid_mats = cnhip.fluid_material(0.03)
    oo = oracle.OracleScene(off)
    for _ in range(5):
        oracle.step(oo, params(cap=64))
    assert not np.array_equal(oo.vel, osc.vel)


def test_vorticity_mixed_materials_bitexact(solver, oracle):
    sc = swirl(mixed_scene())
    r = 0.02
    sc.fluid_mats = np.concatenate([cnhip.fluid_material(r, vorticity_confinement=0.8),
                                    cnhip.fluid_material(r, viscosity=50.0, cohesion=0.5)])
    res, _ = step_pair(sc, params(cap=64), 5, solver, oracle)
    print(res)
    assert all(res), res


def test_vorticity_truncated_lists_bitexact(solver, oracle):
    sc = swirl(scenes.dam_break_64k(side=10))
    sc.fluid_mats = cnhip.fluid_material(0.03, vorticity_confinement=0.6)
    res, _ = step_pair(sc, params(cap=9, substeps=2, iters=2), 4, solver, oracle)
    print(res)
    assert all(res), res


@pytest.mark.parametrize("builder", ["dam", "mixed"])
def test_render_anisotropy_bitexact(solver, oracle, builder):
    """physics_fluid_build_render_anisotropy.cs.hlsl after the step: frames and scales bit for bit,
    the identity frame for solid and pinned particles; unreadable before a step writes them."""
    if builder == "dam":
        sc = swirl(scenes.dam_break_64k(side=12))
    else:
        sc = swirl(mixed_scene())
        sc.fluid_mats = np.concatenate([cnhip.fluid_material(0.02, vorticity_confinement=0.8),
                                        cnhip.fluid_material(0.02, viscosity=50.0, cohesion=0.5)])
    p = params(cap=64, substeps=2, iters=3)
    osc = oracle.OracleScene(sc)
    want = osc.enable_render_anisotropy()
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    solver.render_anisotropy_enable(True)
    with pytest.raises(cnhip.CnError):
        solver.render_anisotropy()
    for _ in range(3):
        solver.step(p)
        oracle.step(osc, p)
        got = solver.render_anisotropy()
        assert np.array_equal(got, want), np.abs(got - want).max()
    fixed = (sc.kinds != cnhip.CN_PARTICLE_KIND_FLUID) | (sc.pos[:, 3] <= 1e-6)
    if fixed.any():
        assert np.array_equal(got[:, fixed, :3], np.broadcast_to(np.eye(3, dtype=np.float32)[:, None, :],
                                                                (3, int(fixed.sum()), 3)))
    pos, _, vel = solver.readback()
    assert np.array_equal(pos, osc.pos) and np.array_equal(vel, osc.vel)
    solver.render_anisotropy_enable(False)
