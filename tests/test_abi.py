"""C-ABI checks that need no GPU: the library loads, exports every declared symbol, and its host
authoring helpers agree bit for bit with the oracle's restatement."""
import ctypes as C
import os
import re

import numpy as np

import cnhip
from cnhip import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_\s\*]*?\b(cn_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = cnhip.load_library()
    declared = _declared_functions()
    assert len(declared) >= 20
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding table covers the same set
    assert declared == set(abi.SIGNATURES), declared ^ set(abi.SIGNATURES)
    assert lib.cn_abi_version() == 1


def test_default_params_match_reference_defaults(oracle):
    p = cnhip.default_params()
    q = abi.StepParams()
    oracle.load().cno_default_params(C.byref(q))
    assert list(p.solver_config) == list(q.solver_config)
    assert p.substeps == 1 and p.default_iterations == 20 and p.max_fluid_neighbors == 128
    assert tuple(p.gravity)[:3] == (0.0, np.float32(-9.81), 0.0)
    # physics_solver.h:101-112 fluid settings
    assert np.float32(p.solver_config[32]) == np.float32(0.7) and np.float32(p.solver_config[35]) == np.float32(0.08)


def test_fluid_material_bitwise_equals_oracle(oracle):
    for r in (0.03, 0.0086022, 0.01, 0.2, 1e-5):
        a = cnhip.fluid_material(r, viscosity=0.02, cohesion=0.3, surface_tension=0.1, cfl_coefficient=0.5)
        b = oracle.fluid_material(r, viscosity=0.02, cohesion=0.3, surface_tension=0.1, cfl=0.5)
        np.testing.assert_array_equal(a.view(np.float32), b)


def test_grid_cell_size_equals_oracle(oracle):
    lib = cnhip.load_library()
    rng = np.random.default_rng(0)
    radii = rng.uniform(0.01, 0.05, 100).astype(np.float32)
    fm = np.concatenate([cnhip.fluid_material(0.03), cnhip.fluid_material(0.02)])
    idx = rng.integers(0, 2, 100).astype(np.uint32)
    idx[::3] = 0xFFFFFFFF
    a = lib.cn_particle_grid_cell_size(100, radii.ctypes.data, idx.ctypes.data, 2, fm.ctypes.data)
    b = oracle.load().cno_grid_cell_size(100, radii.ctypes.data, idx.ctypes.data, 2, fm.ctypes.data)
    assert a == b


def test_struct_layouts_match_header():
    assert C.sizeof(abi.BoxCollider) == 64
    assert C.sizeof(abi.FluidMaterial) == 48
    assert C.sizeof(abi.StepParams) == 4 * (6 + 4 + 40 + 4)
    assert abi.Meta.fluid_neighbor_total.offset == 16
    assert abi.Meta.contact_overflows.offset == 12
    assert C.sizeof(abi.TuneResult) == 24
    assert C.sizeof(abi.Meta) == 48 and abi.Meta.candidate_reuses.offset == 40
