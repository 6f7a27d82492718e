"""The C++ host mirror (cressim-neo_amd/host): the C++ tests shaped like the reference's own
tests/physics executables, each a program that returns 0 on success."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")


def run(name):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), f"{exe} is not built (__graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    return r.stdout


def test_world_soa_layout():
    # CPU only: authoring and the record layouts (the library is loaded, no device is touched)
    if not os.path.exists(os.path.join(BIN, "physics_world_soa_layout")):
        import sys
        sys.path.insert(0, ROOT)
        import __graft_entry__
        __graft_entry__.build_host_tests()
    assert "checks passed" in run("physics_world_soa_layout")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["physics_fluid_compute_smoke", "physics_solver_broad_phase_overflow",
                                  "physics_soft_gpu_scene_upload", "physics_world_revision_upload"])
def test_cpp_gpu(name):
    print(f"running {name}", flush=True)
    assert "checks passed" in run(name)
