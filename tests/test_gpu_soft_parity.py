"""GPU parity of the XPBD soft-body / cloth passes and of mixed scenes against the CPU oracle."""
import numpy as np
import pytest

import cnhip
from cnhip import scenes

pytestmark = pytest.mark.gpu


def params(iters=4, substeps=4, **kw):
    return cnhip.default_params(dt=1 / 120, substeps=substeps, default_iterations=iters, max_fluid_neighbors=64,
                                **kw)


def run_pair(sc, p, steps, solver, oracle):
    osc = oracle.OracleScene(sc)
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    out = []
    for _ in range(steps):
        solver.step(p)
        oracle.step(osc, p)
        pos, _, vel = solver.readback()
        assert np.isfinite(pos).all()
        out.append((float(np.abs(pos - osc.pos).max()), float(np.abs(vel - osc.vel).max()),
                    bool((pos == osc.pos).all() and (vel == osc.vel).all())))
    return out


def soft_and_cloth():
    return scenes.compose([scenes.soft_block_part(6, 0.05), scenes.cloth_part(16, 0.03)], cnhip.fluid_material(0.03),
                          scenes.container((0, 0, 0), (4, 2, 2)), "soft_cloth")


@pytest.mark.parametrize("iters", [1, 4, 7])
def test_soft_and_cloth_bitexact(solver, oracle, iters):
    res = run_pair(soft_and_cloth(), params(iters=iters), 5, solver, oracle)
    print(res)
    assert all(b for _, _, b in res), res


def test_constraint_fields_edge_cases(solver, oracle):
    """Disabled edges, negative and NaN compliances and a negative tet rest volume: the device
    uploads compact constraint forms (compliance and rest volume clamped at zero, a disabled edge
    as a NaN compliance), which must step as the C-ABI records do in the oracle."""
    sc = soft_and_cloth()
    rng = np.random.default_rng(12)
    e = sc.edges
    e["enabled"][rng.random(len(e)) < 0.2] = 0
    e["compliance"][rng.random(len(e)) < 0.1] = -1e-3
    e["compliance"][3] = np.nan
    e["enabled"][3] = 1
    t = sc.tets
    t["compliance"][rng.random(len(t)) < 0.1] = -5e-4
    t["rest_volume"][rng.random(len(t)) < 0.05] *= -1.0
    d = sc.dihedrals
    d["compliance"][rng.random(len(d)) < 0.1] = -2e-3
    res = run_pair(sc, params(), 4, solver, oracle)
    print(res)
    assert all(b for _, _, b in res), res


def test_soft_iterations_differ_from_fluid_iterations(solver, oracle):
    sc = scenes.mixed(12, 6, 16)
    p = params(iters=3, fluid_iterations=2, soft_internal_iterations=5)
    res = run_pair(sc, p, 3, solver, oracle)
    assert all(b for _, _, b in res), res


def test_mixed_small_bitexact(solver, oracle):
    res = run_pair(scenes.mixed(16, 8, 32), params(), 6, solver, oracle)
    print(res)
    assert all(b for _, _, b in res), res


def test_mixed_full_size_one_step(solver, oracle):
    sc = scenes.mixed()  # 262,144 fluid + 32^3 tet grid (148,955 tets) + 256 x 256 cloth
    assert len(sc.tets) == 148_955 and len(sc.dihedrals) == 194_565
    res = run_pair(sc, params(), 1, solver, oracle)
    print(res)
    assert all(b for _, _, b in res), res
