"""Scene edits without a re-upload: colliders (cn_scene_set_colliders: moved, resized, disabled
boxes), material tables (cn_scene_set_materials) and constraint parameters
(cn_scene_set_constraints).  After each edit the steps match the oracle stepping the edited scene
bit for bit, in a single scene and in scenes of a pool; changed counts, a new smoothing radius and
constraints on other particles are refused."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

from test_gpu_fluid_parity import params, small_cloud

pytestmark = pytest.mark.gpu


def edited(cols):
    new = scenes.container((0.0, 0.0, 0.0), (0.6, 0.42, 0.6))   # a lower lid
    assert len(new) == len(cols)
    new["center"][0] += np.float32(0.03)                         # the x-low wall moved inward
    new["enabled"][5] = 0                                        # the z-high wall removed
    return new


def test_set_colliders_single_scene_bitexact(solver, oracle):
    sc = small_cloud(1500, seed=7)
    p = params(cap=64, substeps=2, iters=3)
    osc = oracle.OracleScene(sc)
    keep = sc.copy()
    keep.colliders = sc.colliders.copy()                         # OracleScene aliases the array
    old = oracle.OracleScene(keep)                               # keeps the uploaded colliders
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    for _ in range(2):
        solver.step(p)
        oracle.step(osc, p)
        oracle.step(old, p)
    new = edited(sc.colliders)
    solver.set_scene_colliders(0, new)
    osc.colliders[:] = new
    for s in range(4):
        solver.step(p)
        oracle.step(osc, p)
        pos, _, vel = solver.readback()
        assert np.array_equal(pos, osc.pos) and np.array_equal(vel, osc.vel), s
        oracle.step(old, p)
    assert not np.array_equal(old.pos, osc.pos)                 # the edit took effect
    with pytest.raises(cnhip.CnError):
        solver.set_scene_colliders(0, new[:-1])


def test_set_colliders_in_pool_bitexact(oracle):
    ses = [small_cloud(900, seed=11), small_cloud(700, seed=12)]
    p = params(cap=64, substeps=2, iters=2)
    oscs = [oracle.OracleScene(sc) for sc in ses]
    pool = cnhip.Solver(0)
    try:
        pool.reserve(sum(sc.n for sc in ses), 64)
        pool.upload_batch(ses)
        pool.step(p)
        for o in oscs:
            oracle.step(o, p)
        new = edited(ses[1].colliders)
        pool.set_scene_colliders(1, new)
        oscs[1].colliders[:] = new
        for _ in range(3):
            pool.step(p)
            for o in oscs:
                oracle.step(o, p)
        for k, o in enumerate(oscs):
            pos, _, vel = pool.readback_scene(k)
            assert np.array_equal(pos, o.pos) and np.array_equal(vel, o.vel), k
    finally:
        pool.close()


def test_set_materials_bitexact(solver, oracle):
    """cn_scene_set_materials: new viscosity, cohesion, surface tension (switching the tension
    passes on), vorticity and particle damping take effect on the next step, bit for bit with the
    oracle stepping the edited tables; a new smoothing radius is refused."""
    from test_gpu_fluid_parity import mixed_scene
    sc = mixed_scene(2500, seed=31)
    sc.fluid_mats = sc.fluid_mats.copy()
    sc.part_mats = sc.part_mats.copy()
    p = params(cap=64, substeps=2, iters=3)
    osc = oracle.OracleScene(sc)
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    for _ in range(2):
        solver.step(p)
        oracle.step(osc, p)
    r = 0.02
    fm = np.concatenate([cnhip.fluid_material(r, viscosity=5.0, vorticity_confinement=0.7),
                         cnhip.fluid_material(r, viscosity=20.0, cohesion=0.9, surface_tension=0.4)])
    pm = np.array([[0.1, 0.0, 0.3, -1.0]], np.float32)
    solver.set_scene_materials(0, fm, pm)
    osc.fluid_mats[:] = fm
    osc.part_mats[:] = pm
    for s in range(3):
        solver.step(p)
        oracle.step(osc, p)
        pos, _, vel = solver.readback()
        assert np.array_equal(pos, osc.pos) and np.array_equal(vel, osc.vel), s
    bad = fm.copy()
    bad["smoothing_radius"][0] *= np.float32(1.5)
    with pytest.raises(cnhip.CnError):
        solver.set_scene_materials(0, bad)
    with pytest.raises(cnhip.CnError):
        solver.set_scene_materials(0, fm[:1])


def test_set_constraints_bitexact(oracle):
    """cn_scene_set_constraints: new compliances, rest values and edge enable flags of a tet block
    and a cloth in a pool take effect on the next step, bit for bit with the oracle stepping the
    edited constraints; other particles or other counts are refused."""
    ses = [scenes.batch_scene(1500, 1, seed=41), scenes.batch_scene(1200, 2, seed=42)]
    p = params(cap=64, substeps=2, iters=3)
    oscs = [oracle.OracleScene(sc) for sc in ses]
    pool = cnhip.Solver(0)
    try:
        pool.reserve(sum(sc.n for sc in ses), 64)
        pool.upload_batch(ses)
        for _ in range(2):
            pool.step(p)
            for o in oscs:
                oracle.step(o, p)
        eds = [sc.copy() for sc in ses]
        for ed in eds:
            if ed.edges is not None and len(ed.edges):
                ed.edges["rest_length"] *= np.float32(1.05)
                ed.edges["compliance"] = np.float32(2e-5)
                ed.edges["enabled"][::7] = 0
            if ed.tets is not None and len(ed.tets):
                ed.tets["rest_volume"] *= np.float32(0.9)
                ed.tets["compliance"] = np.float32(1e-5)
            if ed.bends is not None and len(ed.bends):
                ed.bends["compliance"] = np.float32(3e-3)
            if ed.dihedrals is not None and len(ed.dihedrals):
                ed.dihedrals["rest_angle"] += np.float32(0.1)
                ed.dihedrals["compliance"] = np.float32(1e-3)
        for k, (ed, o) in enumerate(zip(eds, oscs)):
            pool.set_scene_constraints(k, ed)
            e, b, d, t = ed.constraint_arrays()
            o.edges[:] = e
            o.bends[:] = b
            o.dihedrals[:] = d
            o.tets[:] = t
        for s in range(3):
            pool.step(p)
            for o in oscs:
                oracle.step(o, p)
            for k, o in enumerate(oscs):
                pos, _, vel = pool.readback_scene(k)
                assert np.array_equal(pos, o.pos) and np.array_equal(vel, o.vel), (s, k)
        bad = eds[0].copy()
        if bad.edges is not None and len(bad.edges):
            bad.edges["b"][0] = bad.edges["a"][0]
        else:
            bad.tets["p"][0, 0] = bad.tets["p"][0, 1]
        with pytest.raises(cnhip.CnError):
            pool.set_scene_constraints(0, bad)
        with pytest.raises(cnhip.CnError):
            pool.set_scene_constraints(0, eds[1])
    finally:
        pool.close()


def vorticity_launches(solver, p) -> int:
    """Vorticity-pass launches of one profiled step (the context's statistics accumulate)."""
    before = solver.profile_read().get("fluid_vorticity", (0, 0.0))[0]
    solver.profile(True)
    solver.step(p)
    after = solver.profile_read().get("fluid_vorticity", (0, 0.0))[0]
    solver.profile(False)
    return after - before


def test_set_materials_unused_vorticity_runs_no_vorticity_pass(solver, oracle):
    """A table edit that gives vorticity confinement only to a material no fluid particle
    references leaves the vorticity passes off (as an upload of that table would) and the steps
    equal the oracle's."""
    sc = small_cloud(1200, seed=17)
    r = float(sc.radii[0])
    sc.fluid_mats = np.concatenate([sc.fluid_mats, cnhip.fluid_material(r)])   # material 1: unused
    p = params(cap=64, substeps=2, iters=2)
    osc = oracle.OracleScene(sc)
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    solver.step(p)
    oracle.step(osc, p)
    fm = sc.fluid_mats.copy()
    fm[1] = cnhip.fluid_material(r, vorticity_confinement=0.8)[0]
    solver.set_scene_materials(0, fm)
    osc.fluid_mats[:] = fm
    assert vorticity_launches(solver, p) == 0
    oracle.step(osc, p)
    pos, _, vel = solver.readback()
    assert np.array_equal(pos, osc.pos) and np.array_equal(vel, osc.vel)
    fm[0] = cnhip.fluid_material(r, vorticity_confinement=0.8)[0]   # now a used material confines
    solver.set_scene_materials(0, fm)
    osc.fluid_mats[:] = fm
    assert vorticity_launches(solver, p) > 0
    oracle.step(osc, p)
    pos, _, vel = solver.readback()
    assert np.array_equal(pos, osc.pos) and np.array_equal(vel, osc.vel)
