"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py from the oracle).

CPU: the oracle still reproduces every committed vector bit for bit.
GPU: libcn_hip.so reproduces them bit for bit through the C ABI.
"""
import os

import numpy as np
import pytest

import cnhip
from cnhip import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def params(**kw):
    return cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64, **kw)


def dam_scene():
    return scenes.dam_break_64k(side=10)


def soft_scene():
    return scenes.compose([scenes.soft_block_part(5, 0.05), scenes.cloth_part(12, 0.03),
                           scenes.fluid_part(6, 0.05, 0.03, origin=(1.0, 0.05, 1.0), seed=3, jitter=0.005)],
                          cnhip.fluid_material(0.03), scenes.container((0, 0, 0), (4, 2, 2)), "golden_soft")


# ------------------------------------------------------------------------------------------ CPU
def test_oracle_reproduces_broadphase_vectors(oracle):
    g = load("broadphase_600.npz")
    keys, cells = oracle.broadphase_entries(g["pos"], float(g["cell"]))
    np.testing.assert_array_equal(keys, g["keys"])
    np.testing.assert_array_equal(cells, g["cells"])
    sk, si = oracle.radix_sort_1bit(keys, np.arange(len(keys), dtype=np.uint32))
    np.testing.assert_array_equal(sk, g["sorted_keys"])
    np.testing.assert_array_equal(si, g["sorted_idx"])
    np.testing.assert_array_equal(oracle.cell_ranges(sk, g["table"].shape[0]), g["table"])


@pytest.mark.parametrize("name,build,steps", [("dam_break_1000.npz", dam_scene, 3),
                                              ("soft_cloth_fluid.npz", soft_scene, 2)])
def test_oracle_reproduces_step_vectors(oracle, name, build, steps):
    g = load(name)
    sc = build()
    np.testing.assert_array_equal(sc.pos, g["pos0"])  # the scene builder has not drifted
    osc = oracle.OracleScene(sc)
    for s in range(steps):
        oracle.step(osc, params())
        np.testing.assert_array_equal(osc.pos, g[f"pos{s + 1}"])
        np.testing.assert_array_equal(osc.vel, g[f"vel{s + 1}"])


def test_oracle_reproduces_neighbor_vectors(oracle):
    g = load("dam_break_1000.npz")
    sc = dam_scene()
    nb = oracle.substep_neighbors(oracle.OracleScene(sc), params(), float(np.max(sc.fluid_mats["smoothing_radius"])))
    assert int(g["counts"].max()) <= g["lists"].shape[1]
    np.testing.assert_array_equal(nb["counts"], g["counts"])
    for i, c in enumerate(g["counts"]):
        np.testing.assert_array_equal(nb["lists"][i, :c], g["lists"][i, :c])


# ------------------------------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_gpu_broadphase_vectors(solver):
    g = load("broadphase_600.npz")
    keys, cells = solver.pass_keys(g["pos"], float(g["cell"]))
    np.testing.assert_array_equal(keys, g["keys"])
    np.testing.assert_array_equal(cells, g["cells"])
    sk, si = solver.pass_sort(keys, np.arange(len(keys), dtype=np.uint32))
    np.testing.assert_array_equal(sk, g["sorted_keys"])
    np.testing.assert_array_equal(si, g["sorted_idx"])
    np.testing.assert_array_equal(solver.pass_cell_ranges(sk, g["table"].shape[0]), g["table"])


@pytest.mark.gpu
def test_gpu_neighbor_vectors(solver):
    g = load("dam_break_1000.npz")
    sc = dam_scene()
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    _, _, counts, lists = solver.pass_substep_broadphase(params(key_mode=cnhip.CN_KEY_HASH))
    np.testing.assert_array_equal(counts, g["counts"])
    for i, c in enumerate(g["counts"]):
        np.testing.assert_array_equal(lists[i, :c], g["lists"][i, :c])


@pytest.mark.gpu
@pytest.mark.parametrize("name,build,steps", [("dam_break_1000.npz", dam_scene, 3),
                                              ("soft_cloth_fluid.npz", soft_scene, 2)])
def test_gpu_step_vectors(solver, name, build, steps):
    g = load(name)
    sc = build()
    solver.reserve(sc.n, 64)
    solver.upload(sc)
    for s in range(steps):
        solver.step(params())
        pos, _, vel = solver.readback()
        np.testing.assert_array_equal(pos, g[f"pos{s + 1}"])
        np.testing.assert_array_equal(vel, g[f"vel{s + 1}"])
