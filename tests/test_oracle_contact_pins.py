"""Pins of the oracle's particle-particle contacts (CPU only): closed forms of one Jacobi contact
solve between two free particles, the candidate filters of IsValidParticleParticleCandidate
(self collision inside a phase group and its constraint-adjacency filter included), the velocity
solve and the partner-list capacity."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

F = np.float32
SLOP, RELAX, MAX_CORR = 1.0e-3, 0.95, 0.05   # solver_config[0], [2], [3] defaults


def _pair(p0, p1, w=(1.0, 1.0), r=(0.01, 0.01), v=((0, 0, 0), (0, 0, 0)), same_part=False, kinds=(0, 0),
          env=(0, 0), part_mat=None):
    parts = []
    for k in range(2):
        part = scenes._solid_part(np.array([p0 if k == 0 else p1], np.float32), r[k], 1.0)
        part["pos"][:, 3] = w[k]
        part["kinds"][:] = kinds[k]
        part["env"][:] = env[k]
        if kinds[k] == 1:
            part["fluid_mat"][:] = 0
        parts.append(part)
    if same_part:
        parts = [{key: np.concatenate([parts[0][key], parts[1][key]]) for key in parts[0]}]
    sc = scenes.compose(parts, cnhip.fluid_material(0.03), np.zeros(0, cnhip.COLLIDER_DTYPE), "pair")
    sc.vel[:, :3] = np.asarray(v, np.float32)
    if part_mat is not None:
        sc.part_mats = np.array([part_mat], np.float32)
    return sc


def _params(iters=1, substeps=1):
    return cnhip.default_params(dt=1 / 120, substeps=substeps, default_iterations=iters, max_fluid_neighbors=64,
                                gravity=(0.0, 0.0, 0.0, 0.0))


def _run(oracle, sc, p=None):
    osc = oracle.OracleScene(sc)
    st = oracle.step(osc, p or _params())
    return osc, st


def test_overlapping_pair_separates_by_the_closed_form(oracle):
    # radii 0.01 + 0.01 at distance 0.015: penetration 0.005, minus the slop, split by inverse mass
    # and scaled by the contact relaxation; the velocity update then carries the motion / dt
    osc, _ = _run(oracle, _pair((0, 0, 0), (0.015, 0, 0)))
    lam = (0.005 - SLOP) / 2 * RELAX
    assert osc.pos[0, 0] == pytest.approx(-lam, abs=2e-8)
    assert osc.pos[1, 0] == pytest.approx(0.015 + lam, abs=2e-8)
    assert osc.pos[0, 1] == 0 and osc.pos[0, 2] == 0
    assert osc.vel[0, 0] == pytest.approx(-lam * 120, rel=1e-5)
    # unequal inverse masses: displacements in the ratio of the inverse masses
    osc, _ = _run(oracle, _pair((0, 0, 0), (0.015, 0, 0), w=(1.0, 3.0)))
    lam = (0.005 - SLOP) / 4 * RELAX
    assert osc.pos[0, 0] == pytest.approx(-lam, abs=2e-8)
    assert osc.pos[1, 0] - F(0.015) == pytest.approx(3 * lam, abs=2e-8)
    # a pinned particle takes nothing and its partner takes its own share only
    osc, _ = _run(oracle, _pair((0, 0, 0), (0.015, 0, 0), w=(0.0, 1.0)))
    assert osc.pos[0, 0] == 0
    assert osc.pos[1, 0] - F(0.015) == pytest.approx((0.005 - SLOP) * RELAX, abs=2e-8)


def test_correction_is_capped_and_coincident_pair_uses_up(oracle):
    # deep overlap: the correction is capped at soft max correction before the mass split
    osc, _ = _run(oracle, _pair((0, 0, 0), (0.02, 0, 0), r=(0.1, 0.1)))
    lam = MAX_CORR / 2 * RELAX
    assert osc.pos[0, 0] == pytest.approx(-lam, abs=1e-7)
    # coincident particles: normal (0, 1, 0) from the lower index, penetration = radius sum
    osc, _ = _run(oracle, _pair((0.3, 0.2, 0.1), (0.3, 0.2, 0.1)))
    lam = (0.02 - SLOP) / 2 * RELAX
    assert osc.pos[0, 1] == pytest.approx(0.2 - lam, abs=1e-7)
    assert osc.pos[1, 1] == pytest.approx(0.2 + lam, abs=1e-7)
    assert osc.pos[0, 0] == F(0.3) and osc.pos[1, 2] == F(0.1)


@pytest.mark.parametrize("case", ["same_group", "fluid_pair", "other_env", "apart", "both_pinned"])
def test_candidate_filters(oracle, case):
    kw = dict(same_part=case == "same_group", kinds=(1, 1) if case == "fluid_pair" else (0, 0),
              env=(0, 1) if case == "other_env" else (0, 0), w=(0.0, 0.0) if case == "both_pinned" else (1.0, 1.0))
    p1 = (0.0201, 0, 0) if case == "apart" else (0.015, 0, 0)
    sc = _pair((0, 0, 0), p1, **kw)
    if case == "fluid_pair":
        sc.pos[:, 3] = 1.0
        # fluid pairs are left to the density constraint: compare with the same pair far apart
        far = _pair((0, 0, 0), (5.0, 0, 0), **kw)
        far.pos[:, 3] = 1.0
        a, _ = _run(oracle, sc, _params())
        b, _ = _run(oracle, far, _params())
        assert a.pos[0, 0] == b.pos[0, 0]
        return
    osc, _ = _run(oracle, sc)
    np.testing.assert_array_equal(osc.pos[:, :3], sc.pos[:, :3])


def test_layer_mask_filter(oracle):
    sc = _pair((0, 0, 0), (0.015, 0, 0))
    sc.layer[1] = 2
    sc.mask[0] = 1  # particle 0 does not collide with layer 2
    osc, _ = _run(oracle, sc)
    np.testing.assert_array_equal(osc.pos[:, :3], sc.pos[:, :3])


def test_static_friction_cancels_tangential_slip(oracle):
    # b slides 0.001 m along z in the substep and ends level with a (normal exactly x); with static
    # friction 1 the whole tangential displacement (<= static * penetration) is removed, relaxed
    # and split by inverse mass
    slip = F(0.12) * F(1 / 120)
    sc = _pair((0, 0, 0), (0.015, 0, -slip), v=((0, 0, 0), (0, 0, 0.12)), part_mat=(1.0, 0.0, 0.0, 1.0))
    osc, _ = _run(oracle, sc)
    tl = slip / 2 * RELAX
    assert osc.pos[0, 2] == pytest.approx(tl, abs=1e-9)
    assert osc.pos[1, 2] == pytest.approx(-tl, abs=1e-9)
    assert osc.pos[0, 0] == pytest.approx(-(0.005 - SLOP) / 2 * RELAX, abs=2e-8)
    # without friction (the default contact material) the slip stays
    osc, _ = _run(oracle, _pair((0, 0, 0), (0.015, 0, -slip), v=((0, 0, 0), (0, 0, 0.12))))
    assert osc.pos[0, 2] == 0 and abs(float(osc.pos[1, 2])) < 1e-9


def test_velocity_solve_removes_approach(oracle):
    # two particles running into each other: after the position solve they still approach, and
    # one velocity iteration (no restitution: penetration above 2 * slop) zeroes the normal
    # relative velocity
    sc = _pair((0, 0, 0), (0.025, 0, 0), v=((1, 0, 0), (-1, 0, 0)))
    osc, _ = _run(oracle, sc)
    dt = F(1 / 120)
    pa, pb = F(0) + F(1) * dt, F(0.025) - F(1) * dt
    pen = F(0.02) - (pb - pa)
    lam = (pen - SLOP) / 2 * RELAX
    assert osc.pos[0, 0] == pytest.approx(pa - lam, abs=1e-7)
    assert abs(float(osc.vel[1, 0] - osc.vel[0, 0])) < 1e-5
    assert abs(float(osc.vel[0, 0] + osc.vel[1, 0])) < 1e-5
    # separating particles keep their velocities
    sc = _pair((0, 0, 0), (0.015, 0, 0), v=((-1, 0, 0), (1, 0, 0)))
    osc, _ = _run(oracle, sc)
    assert osc.vel[1, 0] - osc.vel[0, 0] > 1.9


def test_partner_overflow_is_counted(oracle):
    # one particle overlapping 40 particles of another group: 8 more candidates than a list holds
    rng = np.random.default_rng(0)
    cluster = (rng.uniform(-0.004, 0.004, (40, 3))).astype(np.float32)
    a = scenes._solid_part(np.zeros((1, 3), np.float32), 0.01, 1.0)
    b = scenes._solid_part(cluster, 0.01, 1.0)
    sc = scenes.compose([a, b], cnhip.fluid_material(0.03), np.zeros(0, cnhip.COLLIDER_DTYPE), "overflow")
    sc.pos[:, 3] = 0.0  # pinned: candidates do not depend on mass, and nothing moves between substeps
    _, st = _run(oracle, sc)
    assert st.contact_overflows == 1
    _, st = _run(oracle, sc, _params(substeps=3))
    assert st.contact_overflows == 3


SELF = np.uint32(0x80000000)


def _one_group(flags=(True, True), adjacent=False):
    """Two overlapping solid particles of one phase group (radii 0.01 + 0.01 at distance 0.015),
    each with its self-collide flag as given, constraint-adjacent or not."""
    sc = _pair((0, 0, 0), (0.015, 0, 0), same_part=True)
    assert sc.phase[0] == sc.phase[1]
    for k in range(2):
        if flags[k]:
            sc.phase[k] |= SELF
    if adjacent:
        from cnhip import authoring as au
        sc.adjacency = au.adjacency(2, [(0, 1)])
    else:
        sc.adjacency = (np.zeros(2, np.uint32), np.zeros(2, np.uint32), np.zeros(0, np.uint32))
    return sc


def test_self_collision_inside_a_group(oracle):
    """IsValidParticleParticleCandidate's group branch: one group collides with itself only when
    both particles have the self-collide flag and they are not constraint-adjacent; then the pair
    separates exactly as a pair of two groups does."""
    lam = (0.005 - SLOP) / 2 * RELAX
    osc, _ = _run(oracle, _one_group((True, True)))
    assert osc.pos[0, 0] == pytest.approx(-lam, abs=2e-8)
    assert osc.pos[1, 0] == pytest.approx(0.015 + lam, abs=2e-8)
    two, _ = _run(oracle, _pair((0, 0, 0), (0.015, 0, 0)))
    np.testing.assert_array_equal(osc.pos, two.pos)
    for flags in ((True, False), (False, True), (False, False)):
        osc, _ = _run(oracle, _one_group(flags))
        assert osc.pos[0, 0] == 0 and osc.pos[1, 0] == F(0.015), flags
    osc, _ = _run(oracle, _one_group((True, True), adjacent=True))
    assert osc.pos[0, 0] == 0 and osc.pos[1, 0] == F(0.015)


def test_adjacency_is_the_reference_construction():
    """Tet soft bodies: the tet edges; cloths: structural edges plus the opposite vertices of
    every bending hinge; both directions, each list sorted and unique
    (soft_body_authoring.cpp:392-408, cloth_authoring.cpp:139-188)."""
    from cnhip import authoring as au
    quad = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [1, 0, 1]], np.float32)
    e, d = au.cloth_constraints(quad, np.array([[0, 1, 3], [0, 3, 2]], np.uint32))
    off, cnt, idx = au.adjacency(4, au.cloth_adjacency_pairs(e, d))
    lists = [idx[off[i]:off[i] + cnt[i]].tolist() for i in range(4)]
    # five structural edges (0-1, 0-2, 0-3, 1-3, 2-3) and the hinge 0-3 whose opposites are 1 and 2
    assert lists == [[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]]
    pos, tets = au.tet_grid((1.0, 1.0, 1.0), resolution=(2, 2, 2))
    edges, _ = au.tet_constraints(pos, tets)
    off, cnt, idx = au.adjacency(len(pos), au.tet_adjacency_pairs(edges))
    want = [set() for _ in range(len(pos))]
    for t in tets:
        for a in range(4):
            for b in range(4):
                if a != b:
                    want[int(t[a])].add(int(t[b]))
    assert [idx[off[i]:off[i] + cnt[i]].tolist() for i in range(len(pos))] == [sorted(w) for w in want]


def test_contact_stack_has_contacts(oracle):
    # the stack's layers interact: moving the fluid into another environment changes the solids
    sc = scenes.contact_stack(4, 12, 3)
    alone = sc.copy()
    alone.env[alone.kinds == 1] = 1
    p = cnhip.default_params(dt=1 / 120, substeps=2, default_iterations=3, max_fluid_neighbors=64)
    a, st = _run(oracle, sc, p)
    b, _ = _run(oracle, alone, p)
    solid = sc.kinds == 0
    assert st.contact_overflows == 0
    assert not np.array_equal(a.pos[solid], b.pos[solid])
