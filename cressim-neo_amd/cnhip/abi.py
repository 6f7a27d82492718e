"""ctypes mirror of include/cn_hip.h and loader for libcn_hip.so.

The library is looked up next to this package (``cressim-neo_amd/libcn_hip.so``).  A missing or
unloadable library raises ``CnLibraryError``: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libcn_hip.so")

CN_OK = 0
CN_ERR_INVALID_ARGUMENT = 1
CN_ERR_OUT_OF_MEMORY = 2
CN_ERR_DEVICE = 3
CN_ERR_NOT_READY = 4
CN_ERR_CAPACITY = 5
CN_ERR_UNSUPPORTED = 6

CN_PARTICLE_KIND_SOLID = 0
CN_PARTICLE_KIND_FLUID = 1
CN_KEY_MORTON = 0
CN_KEY_HASH = 1
CN_SOLVER_CONFIG_FLOATS = 40
CN_STEP_CANDIDATE_REUSE = 1
CN_NOALLOC_MAX_PAIRS = 1 << 20
CN_SHAPE_SPHERE, CN_SHAPE_BOX, CN_SHAPE_CAPSULE = 0, 1, 2
CN_STATE_DEVICE_POINTERS = 1
CN_KERNEL_SLOTS = 16


class CnLibraryError(RuntimeError):
    pass


class CnError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"cn status {status}: {message}")
        self.status = status


class FluidMaterial(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "rest_density", "inv_rest_density", "smoothing_radius", "density_constraint_scale",
        "viscosity", "cohesion", "cohesion1", "cohesion2", "surface_tension",
        "vorticity_confinement", "gravity_scale", "cfl_radius")]


class FluidMaterialDesc(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "viscosity", "cohesion", "surface_tension", "vorticity_confinement", "gravity_scale",
        "cfl_coefficient")]


class BoxCollider(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("half_extents", C.c_float * 3),
                ("material", C.c_float * 4), ("environment", C.c_uint32),
                ("layer", C.c_uint32), ("mask", C.c_uint32), ("enabled", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


class ParticleSoaView(C.Structure):
    _fields_ = [("count", C.c_uint32),
                ("positions_inv_mass", C.c_void_p), ("velocities", C.c_void_p),
                ("radii", C.c_void_p), ("kinds", C.c_void_p),
                ("fluid_material_indices", C.c_void_p),
                ("particle_material_indices", C.c_void_p),
                ("environment_indices", C.c_void_p), ("phases", C.c_void_p),
                ("collision_layers", C.c_void_p), ("collision_masks", C.c_void_p),
                ("adjacency_offsets", C.c_void_p), ("adjacency_counts", C.c_void_p),
                ("adjacency_index_count", C.c_uint32), ("adjacency_indices", C.c_void_p)]


class SceneTables(C.Structure):
    _fields_ = [("fluid_material_count", C.c_uint32), ("fluid_materials", C.c_void_p),
                ("particle_material_count", C.c_uint32), ("particle_materials", C.c_void_p),
                ("collider_count", C.c_uint32), ("colliders", C.c_void_p),
                ("shape_collider_count", C.c_uint32), ("shape_colliders", C.c_void_p)]


class Collider(C.Structure):
    _fields_ = [("shape", C.c_uint32), ("environment", C.c_uint32), ("layer", C.c_uint32), ("mask", C.c_uint32),
                ("body_position", C.c_float * 4), ("body_orientation", C.c_float * 4),
                ("body_scale", C.c_float * 4), ("local_position", C.c_float * 4),
                ("local_orientation", C.c_float * 4), ("shape_params", C.c_float * 4),
                ("material", C.c_float * 4), ("enabled", C.c_uint32), ("reserved", C.c_uint32 * 3)]


class ConstraintView(C.Structure):
    _fields_ = [("edge_count", C.c_uint32), ("edges", C.c_void_p),
                ("bend_count", C.c_uint32), ("bends", C.c_void_p),
                ("dihedral_count", C.c_uint32), ("dihedrals", C.c_void_p),
                ("tet_count", C.c_uint32), ("tets", C.c_void_p)]


class SceneDesc(C.Structure):
    _fields_ = [("particles", C.POINTER(ParticleSoaView)), ("tables", C.POINTER(SceneTables)),
                ("constraints", C.POINTER(ConstraintView))]


class StepParams(C.Structure):
    _fields_ = [("dt", C.c_float), ("substeps", C.c_uint32), ("default_iterations", C.c_uint32),
                ("fluid_iterations", C.c_uint32), ("soft_internal_iterations", C.c_uint32),
                ("soft_contact_iterations", C.c_uint32), ("gravity", C.c_float * 4),
                ("solver_config", C.c_float * CN_SOLVER_CONFIG_FLOATS),
                ("cell_size", C.c_float), ("max_fluid_neighbors", C.c_uint32),
                ("key_mode", C.c_uint32), ("flags", C.c_uint32)]


class Meta(C.Structure):
    _fields_ = [("max_fluid_neighbor_count", C.c_uint32),
                ("fluid_neighbor_truncations", C.c_uint32),
                ("max_cell_occupancy", C.c_uint32), ("contact_overflows", C.c_uint32),
                ("fluid_neighbor_total", C.c_uint64), ("last_particle_count", C.c_uint32),
                ("lds_tiles", C.c_uint32), ("run_table_tiles", C.c_uint32),
                ("max_displacement2_bits", C.c_uint32), ("candidate_reuses", C.c_uint32),
                ("reserved", C.c_uint32)]


class DeviceViews(C.Structure):
    _fields_ = [("positions_inv_mass", C.c_void_p), ("velocities", C.c_void_p),
                ("count", C.c_uint32), ("stream", C.c_void_p)]


class DeviceDesc(C.Structure):
    _fields_ = [("device_ordinal", C.c_int), ("flags", C.c_uint32)]


class KernelStats(C.Structure):
    _fields_ = [("name", (C.c_char * 32) * CN_KERNEL_SLOTS),
                ("launches", C.c_uint32 * CN_KERNEL_SLOTS),
                ("total_ms", C.c_double * CN_KERNEL_SLOTS), ("used", C.c_uint32)]


class TuneResult(C.Structure):
    _fields_ = [("tile_max", C.c_uint32), ("grid_max", C.c_uint32), ("candidates", C.c_uint32),
                ("reserved", C.c_uint32), ("ms_per_step", C.c_float), ("default_ms_per_step", C.c_float)]


# every symbol include/cn_hip.h declares, with its ctypes signature
_P = C.c_void_p
_U = C.c_uint32
SIGNATURES = {
    "cn_abi_version": (_U, []),
    "cn_ctx_create": (C.c_int, [C.POINTER(DeviceDesc), C.POINTER(_P)]),
    "cn_ctx_destroy": (None, [_P]),
    "cn_last_error": (C.c_char_p, [_P]),
    "cn_default_step_params": (None, [C.POINTER(StepParams)]),
    "cn_fluid_material_from_desc": (C.c_int, [C.POINTER(FluidMaterialDesc), C.c_float,
                                               C.POINTER(FluidMaterial)]),
    "cn_particle_grid_cell_size": (C.c_float, [_U, _P, _P, _U, _P]),
    "cn_scene_reserve": (C.c_int, [_P, _U, _U]),
    "cn_scene_upload": (C.c_int, [_P, C.POINTER(ParticleSoaView), C.POINTER(SceneTables),
                                  C.POINTER(ConstraintView)]),
    "cn_step": (C.c_int, [_P, C.POINTER(StepParams)]),
    "cn_batch_upload": (C.c_int, [_P, _U, C.POINTER(SceneDesc)]),
    "cn_batch_scene_range": (C.c_int, [_P, _U, C.POINTER(_U), C.POINTER(_U)]),
    "cn_graph_enable": (C.c_int, [_P, C.c_int]),
    "cn_tick_policy": (C.c_int, [_P, C.c_float]),
    "cn_concurrent_phases": (C.c_int, [_P, C.c_int]),
    "cn_autotune": (C.c_int, [_P, C.POINTER(StepParams), _U, C.POINTER(TuneResult)]),
    "cn_tune_set": (C.c_int, [_P, _U, _U]),
    "cn_step_scenes": (C.c_int, [_P, C.POINTER(StepParams), _U, _P]),
    "cn_readback": (C.c_int, [_P, _P, _P, _P, _U]),
    "cn_device_views_get": (C.c_int, [_P, C.POINTER(DeviceViews)]),
    "cn_scene_set_state": (C.c_int, [_P, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32]),
    "cn_particles_update_range": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32]),
    "cn_scene_set_colliders": (C.c_int, [_P, C.c_uint32, C.c_void_p, C.c_uint32]),
    "cn_scene_set_materials": (C.c_int, [_P, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]),
    "cn_scene_set_constraints": (C.c_int, [_P, C.c_uint32, C.c_void_p]),
    "cn_render_anisotropy_enable": (C.c_int, [_P, C.c_int]),
    "cn_render_anisotropy_read": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "cn_validate_meta": (C.c_int, [_P, C.POINTER(Meta)]),
    "cn_synchronize": (C.c_int, [_P]),
    "cn_query": (C.c_int, [_P]),
    "cn_pass_broadphase_keys": (C.c_int, [_P, _U, _P, C.c_float, _P, _P]),
    "cn_pass_sort_pairs": (C.c_int, [_P, _U, _P, _P]),
    "cn_pass_sort_pairs_noalloc": (C.c_int, [_P, _U, C.c_void_p, C.c_void_p]),
    "cn_pass_cell_ranges": (C.c_int, [_P, _U, _P, _U, _P]),
    "cn_pass_substep_broadphase": (C.c_int, [_P, C.POINTER(StepParams), _P, _P, _P, _P]),
    "cn_profile_enable": (C.c_int, [_P, C.c_int]),
    "cn_profile_read": (C.c_int, [_P, C.POINTER(KernelStats)]),
}

_lib = None


def torch_hip_runtime() -> str | None:
    """Path of the HIP runtime that torch bundles (``torch/lib/libamdhip64.so``), found without
    importing torch; None when torch is not installed or bundles none."""
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return None
    if spec is None or not spec.origin:
        return None
    p = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    return p if os.path.exists(p) else None


def hip_runtimes_mapped() -> list:
    """The distinct libamdhip64 files mapped into this process (from /proc/self/maps)."""
    seen = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 6 and os.path.basename(parts[5]).startswith("libamdhip64"):
                    seen.add(os.path.realpath(parts[5]))
    except OSError:  # pragma: no cover - not Linux
        return []
    return sorted(seen)


def require_single_hip_runtime(what: str):
    """Raises CnLibraryError when two HIP runtimes are mapped: device pointers and streams of one
    are not valid in the other, so handing the library's buffers to torch would be undefined."""
    rts = hip_runtimes_mapped()
    if len(rts) > 1:
        raise CnLibraryError(
            f"{what}: two HIP runtimes are mapped ({', '.join(rts)}); load libcn_hip.so through "
            "cnhip.load_library() (which maps torch's runtime first) or import torch before loading it")


def load_library(path: str | None = None) -> C.CDLL:
    """Load libcn_hip.so and bind every C-ABI symbol; raises CnLibraryError if absent.

    torch's ROCm build links the unversioned ``libamdhip64.so`` from its own ``lib`` directory,
    while libcn_hip.so needs the SONAME ``libamdhip64.so.7``.  Loaded first, the library would map
    /opt/rocm's runtime, and a later ``import torch`` would map torch's copy as a second, separate
    HIP runtime (its NEEDED name does not match the loaded SONAME).  So when torch is installed its
    runtime is mapped first: its SONAME is ``libamdhip64.so.7``, the library binds to it, and a
    later ``import torch`` finds the same file already loaded.  One runtime serves both, whatever
    the import order."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise CnLibraryError(f"{p} not found: build the HIP library first (__graft_entry__.build())")
    trt = torch_hip_runtime()
    if trt is not None:
        try:
            C.CDLL(trt, mode=C.RTLD_GLOBAL)
        except OSError as e:  # pragma: no cover - a broken torch install
            raise CnLibraryError(f"cannot load torch's HIP runtime {trt}: {e}") from e
    try:
        lib = C.CDLL(p)
    except OSError as e:  # pragma: no cover - depends on the ROCm runtime being present
        raise CnLibraryError(f"cannot load {p}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def ptr(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed across the C ABI must be contiguous"
    return a.ctypes.data
