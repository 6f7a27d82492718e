"""Python front of libcn_hip.so, the gfx950 particle substep behind the CRESSim-Neo solver seam.

``Solver`` mirrors the shape of the reference's ``PhysicsSolver`` (include/physics/physics_solver.h
:150-205): construct with a descriptor, upload a world, ``step()``, read back.  ``Scene`` mirrors the
unified particle pool (``ParticleSoAHost``, include/physics/physics_types.h:858) plus the tables the
solver reads.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .abi import (CN_KEY_HASH, CN_KEY_MORTON, CN_PARTICLE_KIND_FLUID, CN_PARTICLE_KIND_SOLID,
                  CnError, CnLibraryError, load_library)

__all__ = ["Scene", "Solver", "default_params", "fluid_material", "shape_collider", "CnError", "CnLibraryError",
           "CN_KEY_HASH", "CN_KEY_MORTON", "CN_PARTICLE_KIND_FLUID", "CN_PARTICLE_KIND_SOLID",
           "load_library"]

FLUID_MATERIAL_DTYPE = np.dtype([(n, np.float32) for n, _ in abi.FluidMaterial._fields_])
COLLIDER_DTYPE = np.dtype([("center", np.float32, 3), ("half_extents", np.float32, 3),
                           ("material", np.float32, 4), ("environment", np.uint32),
                           ("layer", np.uint32), ("mask", np.uint32), ("enabled", np.uint32),
                           ("reserved", np.uint32, 2)])
assert COLLIDER_DTYPE.itemsize == C.sizeof(abi.BoxCollider) == 64
# cn_collider: a static collider of any shape (sphere, oriented box, capsule) on a static body
SHAPE_COLLIDER_DTYPE = np.dtype([("shape", np.uint32), ("environment", np.uint32), ("layer", np.uint32),
                                 ("mask", np.uint32), ("body_position", np.float32, 4),
                                 ("body_orientation", np.float32, 4), ("body_scale", np.float32, 4),
                                 ("local_position", np.float32, 4), ("local_orientation", np.float32, 4),
                                 ("shape_params", np.float32, 4), ("material", np.float32, 4),
                                 ("enabled", np.uint32), ("reserved", np.uint32, 3)])
assert SHAPE_COLLIDER_DTYPE.itemsize == C.sizeof(abi.Collider) == 144
assert FLUID_MATERIAL_DTYPE.itemsize == C.sizeof(abi.FluidMaterial) == 48


@dataclass
class Scene:
    """Unified particle pool and scene tables (host arrays, particle-index order)."""
    pos: np.ndarray                     # (n, 4) float32 xyz + inverse mass
    vel: np.ndarray                     # (n, 4) float32
    radii: np.ndarray                   # (n,) float32
    kinds: np.ndarray                   # (n,) uint32
    fluid_mat: np.ndarray               # (n,) uint32
    part_mat: np.ndarray                # (n,) uint32
    env: np.ndarray                     # (n,) uint32
    phase: np.ndarray                   # (n,) uint32
    layer: np.ndarray                   # (n,) uint32
    mask: np.ndarray                    # (n,) uint32
    fluid_mats: np.ndarray              # structured FLUID_MATERIAL_DTYPE
    part_mats: np.ndarray               # (m, 4) float32
    colliders: np.ndarray               # structured COLLIDER_DTYPE
    name: str = "scene"
    info: dict = field(default_factory=dict)
    edges: np.ndarray | None = None      # authoring.EDGE_DTYPE
    bends: np.ndarray | None = None      # authoring.BEND_DTYPE
    dihedrals: np.ndarray | None = None  # authoring.DIHEDRAL_DTYPE
    tets: np.ndarray | None = None       # authoring.TET_DTYPE
    # constraint adjacency (offsets (n,), counts (n,), indices) uint32, or None: the particles each
    # particle shares a constraint with, which a self-colliding phase group does not collide with
    adjacency: tuple | None = None
    # static colliders of any shape (SHAPE_COLLIDER_DTYPE), after the boxes of `colliders`
    shape_colliders: np.ndarray | None = None

    def constraint_arrays(self):
        from . import authoring as au
        def arr(a, dt):
            return np.ascontiguousarray(a if a is not None else np.zeros(0, dt))
        return (arr(self.edges, au.EDGE_DTYPE), arr(self.bends, au.BEND_DTYPE),
                arr(self.dihedrals, au.DIHEDRAL_DTYPE), arr(self.tets, au.TET_DTYPE))

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    def meta(self) -> np.ndarray:
        return np.ascontiguousarray(np.stack([self.env, self.phase, self.layer, self.mask], 1)
                                    .astype(np.uint32))

    def copy(self) -> "Scene":
        import copy
        return copy.deepcopy(self)


def default_params(**overrides) -> abi.StepParams:
    """PhysicsSolverDesc defaults (include/physics/physics_solver.h:21-143) as cn_step_params."""
    lib = load_library()
    p = abi.StepParams()
    lib.cn_default_step_params(C.byref(p))
    for k, v in overrides.items():
        if k == "gravity":
            for i, g in enumerate(v):
                p.gravity[i] = g
        else:
            setattr(p, k, v)
    return p


def fluid_material(radius: float, viscosity=0.01, cohesion=0.0, surface_tension=0.0,
                   vorticity_confinement=0.0, gravity_scale=1.0, cfl_coefficient=1.0) -> np.ndarray:
    """toFluidSolverMaterial (src/physics/physics_world.cpp:233-263) through the library."""
    lib = load_library()
    d = abi.FluidMaterialDesc(viscosity, cohesion, surface_tension, vorticity_confinement,
                              gravity_scale, cfl_coefficient)
    out = abi.FluidMaterial()
    st = lib.cn_fluid_material_from_desc(C.byref(d), C.c_float(radius), C.byref(out))
    if st != abi.CN_OK:
        raise CnError(st, "cn_fluid_material_from_desc failed")
    arr = np.zeros(1, FLUID_MATERIAL_DTYPE)
    C.memmove(arr.ctypes.data, C.byref(out), C.sizeof(out))
    return arr


def shape_collider(shape: int, position, orientation=(0.0, 0.0, 0.0, 1.0), params=(0.1, 0.1, 0.1),
                   scale=(1.0, 1.0, 1.0), local_position=(0.0, 0.0, 0.0), local_orientation=(0.0, 0.0, 0.0, 1.0),
                   material=(0.0, 0.0, 0.0, -1.0), environment=0, layer=1, mask=0xFFFFFFFF, enabled=1) -> np.ndarray:
    """One static collider record (SHAPE_COLLIDER_DTYPE): a static body at `position` with
    `orientation` (x, y, z, w) and `scale`, carrying a sphere (params = radius,), an oriented box
    (half extents) or a capsule (radius, half height along local y) at its local pose."""
    out = np.zeros(1, SHAPE_COLLIDER_DTYPE)
    out["shape"] = shape
    out["body_position"][0, :3] = position
    out["body_orientation"][0] = orientation
    out["body_scale"][0, :3] = scale
    out["local_position"][0, :3] = local_position
    out["local_orientation"][0] = local_orientation
    prm = list(params) + [0.0] * (4 - len(params))
    out["shape_params"][0] = prm
    out["material"][0] = material
    out["environment"], out["layer"], out["mask"], out["enabled"] = environment, layer, mask, enabled
    return out


class Solver:
    """One HIP context with one resident scene: PhysicsSolver-shaped driver over the C ABI."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        self._ctx = C.c_void_p()
        desc = abi.DeviceDesc(device, 0)
        st = self.lib.cn_ctx_create(C.byref(desc), C.byref(self._ctx))
        if st != abi.CN_OK:
            raise CnError(st, f"cn_ctx_create failed on device {device}")
        self.scene: Scene | None = None
        self._keep = []
        self.device = device
        self._nviews = 0            # live DLPack exports of the resident buffers
        self._close_pending = False

    # -- helpers
    def _check(self, st: int, what: str):
        if st != abi.CN_OK:
            msg = self.lib.cn_last_error(self._ctx)
            raise CnError(st, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        """Destroys the context; while exported tensor views are alive the buffers they alias are
        kept and the context goes when the last view is released."""
        if self._nviews > 0:
            self._close_pending = True
            return
        if self._ctx:
            self.lib.cn_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- scene
    def reserve(self, n: int, max_fluid_neighbors: int = 128):
        self._refuse_while_viewed("reserve()")
        self._check(self.lib.cn_scene_reserve(self._ctx, n, max_fluid_neighbors), "cn_scene_reserve")

    @staticmethod
    def _views(sc: Scene):
        """C views of a scene's arrays; the returned arrays must outlive the call that uses the views."""
        arrs = dict(
            pos=np.ascontiguousarray(sc.pos, np.float32), vel=np.ascontiguousarray(sc.vel, np.float32),
            radii=np.ascontiguousarray(sc.radii, np.float32),
            kinds=np.ascontiguousarray(sc.kinds, np.uint32),
            fm=np.ascontiguousarray(sc.fluid_mat, np.uint32),
            pm=np.ascontiguousarray(sc.part_mat, np.uint32),
            env=np.ascontiguousarray(sc.env, np.uint32), ph=np.ascontiguousarray(sc.phase, np.uint32),
            ly=np.ascontiguousarray(sc.layer, np.uint32), mk=np.ascontiguousarray(sc.mask, np.uint32),
            fmats=np.ascontiguousarray(sc.fluid_mats), pmats=np.ascontiguousarray(sc.part_mats, np.float32),
            cols=np.ascontiguousarray(sc.colliders))
        p = abi.ptr
        view = abi.ParticleSoaView(sc.n, p(arrs["pos"]), p(arrs["vel"]), p(arrs["radii"]), p(arrs["kinds"]),
                                   p(arrs["fm"]), p(arrs["pm"]), p(arrs["env"]), p(arrs["ph"]),
                                   p(arrs["ly"]), p(arrs["mk"]))
        if sc.adjacency is not None:
            ao, ac, ai = (np.ascontiguousarray(a, np.uint32) for a in sc.adjacency)
            arrs.update(ao=ao, ac=ac, ai=ai)
            view.adjacency_offsets, view.adjacency_counts = p(ao), p(ac)
            view.adjacency_index_count = len(ai)
            view.adjacency_indices = p(ai) if len(ai) else None
        tables = abi.SceneTables(len(arrs["fmats"]), p(arrs["fmats"]), len(arrs["pmats"]), p(arrs["pmats"]),
                                 len(arrs["cols"]), p(arrs["cols"]) if len(arrs["cols"]) else None)
        if sc.shape_colliders is not None and len(sc.shape_colliders):
            arrs["shapes"] = np.ascontiguousarray(sc.shape_colliders)
            assert arrs["shapes"].dtype == SHAPE_COLLIDER_DTYPE
            tables.shape_collider_count = len(arrs["shapes"])
            tables.shape_colliders = p(arrs["shapes"])
        e, b, d, t = sc.constraint_arrays()
        cons = abi.ConstraintView(len(e), p(e) if len(e) else None, len(b), p(b) if len(b) else None,
                                  len(d), p(d) if len(d) else None, len(t), p(t) if len(t) else None)
        return view, tables, cons, (arrs, e, b, d, t)

    def _refuse_while_viewed(self, what: str):
        # an upload may reallocate the buffers exported tensors alias (a device use-after-free),
        # or lay another scene over them
        if self._nviews > 0:
            raise RuntimeError(f"release the exported tensor views before {what} replaces their buffers")

    def upload(self, sc: Scene):
        self._refuse_while_viewed("upload()")
        self.scene = sc
        self.scenes = [sc]
        self.n = sc.n
        view, tables, cons, keep = self._views(sc)
        self._check(self.lib.cn_scene_upload(self._ctx, C.byref(view), C.byref(tables), C.byref(cons)),
                    "cn_scene_upload")

    def upload_batch(self, scenes):
        """Upload independent scenes into one pool (cn_batch_upload); step() then steps them all."""
        self._refuse_while_viewed("upload_batch()")
        self.scene = None
        self.scenes = list(scenes)
        views = [self._views(sc) for sc in self.scenes]
        descs = (abi.SceneDesc * len(views))()
        for k, (v, t, c, _) in enumerate(views):
            descs[k] = abi.SceneDesc(C.pointer(v), C.pointer(t), C.pointer(c))
        self.n = sum(sc.n for sc in self.scenes)
        self._check(self.lib.cn_batch_upload(self._ctx, len(views), descs), "cn_batch_upload")
        self.ranges = []
        for k in range(len(views)):
            first, count = C.c_uint32(), C.c_uint32()
            self._check(self.lib.cn_batch_scene_range(self._ctx, k, C.byref(first), C.byref(count)),
                        "cn_batch_scene_range")
            self.ranges.append((first.value, count.value))

    def readback_scene(self, k: int):
        """(pos, prev, vel) of scene k of the pool."""
        pos, prev, vel = self.readback()
        f, c = self.ranges[k]
        return pos[f:f + c], prev[f:f + c], vel[f:f + c]

    def set_scene_colliders(self, k: int, colliders: np.ndarray):
        """Replace scene k's static box colliders (COLLIDER_DTYPE, the count of its upload)."""
        c = np.ascontiguousarray(colliders, COLLIDER_DTYPE)
        self._check(self.lib.cn_scene_set_colliders(self._ctx, k, c.ctypes.data if len(c) else None, len(c)),
                    "cn_scene_set_colliders")

    def set_scene_materials(self, k: int, fluid_mats=None, particle_mats=None):
        """Replace scene k's fluid materials (FLUID_MATERIAL_DTYPE) and/or particle materials
        (float32 (m, 4)), each with the count of its upload; None leaves a table as it is."""
        f = None if fluid_mats is None else np.ascontiguousarray(fluid_mats, FLUID_MATERIAL_DTYPE)
        pm = None if particle_mats is None else np.ascontiguousarray(particle_mats, np.float32).reshape(-1, 4)
        self._check(self.lib.cn_scene_set_materials(
            self._ctx, k, f.ctypes.data if f is not None and len(f) else None, 0 if f is None else len(f),
            pm.ctypes.data if pm is not None and len(pm) else None, 0 if pm is None else len(pm)),
            "cn_scene_set_materials")

    def set_scene_constraints(self, k: int, sc: Scene):
        """Replace scene k's constraint parameters with those of sc's constraint arrays (same
        counts, order and particles as the upload)."""
        e, b, d, t = sc.constraint_arrays()
        p = abi.ptr
        cons = abi.ConstraintView(len(e), p(e) if len(e) else None, len(b), p(b) if len(b) else None,
                                  len(d), p(d) if len(d) else None, len(t), p(t) if len(t) else None)
        self._check(self.lib.cn_scene_set_constraints(self._ctx, k, C.byref(cons)), "cn_scene_set_constraints")

    def render_anisotropy_enable(self, enable: bool = True):
        """Make every step end with the render anisotropy (cn_render_anisotropy_enable)."""
        self._check(self.lib.cn_render_anisotropy_enable(self._ctx, 1 if enable else 0), "cn_render_anisotropy_enable")

    def render_anisotropy(self) -> np.ndarray:
        """(3, n, 4) float32: axis1, axis2, axis3 of every particle in particle order, w = scale."""
        out = np.zeros((3, self.n, 4), np.float32)
        self._check(self.lib.cn_render_anisotropy_read(self._ctx, out[0].ctypes.data, out[1].ctypes.data,
                                                       out[2].ctypes.data, self.n), "cn_render_anisotropy_read")
        return out

    def step(self, params: abi.StepParams):
        self._check(self.lib.cn_step(self._ctx, C.byref(params)), "cn_step")

    def step_scenes(self, params: abi.StepParams, scenes):
        """Step only the listed scenes of the pool (cn_step_scenes); the others keep their state."""
        ids = np.ascontiguousarray(np.asarray(scenes, np.uint32))
        self._check(self.lib.cn_step_scenes(self._ctx, C.byref(params), len(ids), ids.ctypes.data if len(ids) else None),
                    "cn_step_scenes")

    def tick_policy(self, compact_below: float):
        """cn_tick_policy: a cn_step_scenes whose scenes hold at most this share of the pool's
        particles steps their slots alone (0: always the whole pool)."""
        self._check(self.lib.cn_tick_policy(self._ctx, float(compact_below)), "cn_tick_policy")

    def concurrent_phases(self, enable: bool):
        """cn_concurrent_phases: the soft phase on a second stream beside the fluid passes."""
        self._check(self.lib.cn_concurrent_phases(self._ctx, 1 if enable else 0), "cn_concurrent_phases")

    def graphs(self, enable: bool):
        """Capture cn_step into a HIP graph and replay it (cn_graph_enable)."""
        self._check(self.lib.cn_graph_enable(self._ctx, 1 if enable else 0), "cn_graph_enable")

    def autotune(self, params: abi.StepParams, steps: int = 2) -> dict:
        """Pick the fluid passes' launch shape by timing cn_step from the current state
        (cn_autotune); the state is left as it was."""
        r = abi.TuneResult()
        self._check(self.lib.cn_autotune(self._ctx, C.byref(params), steps, C.byref(r)), "cn_autotune")
        return dict(tile_max=r.tile_max, grid_max=r.grid_max, candidates=r.candidates,
                    ms_per_step=r.ms_per_step, default_ms_per_step=r.default_ms_per_step)

    def tune_set(self, tile_max: int, grid_max: int):
        self._check(self.lib.cn_tune_set(self._ctx, tile_max, grid_max), "cn_tune_set")

    def done(self) -> bool:
        """True when the context's enqueued work has finished (cn_query, non-blocking)."""
        st = self.lib.cn_query(self._ctx)
        if st == abi.CN_ERR_NOT_READY:
            return False
        self._check(st, "cn_query")
        return True

    def synchronize(self):
        self._check(self.lib.cn_synchronize(self._ctx), "cn_synchronize")

    def readback(self):
        n = self.n
        pos = np.zeros((n, 4), np.float32)
        prev = np.zeros((n, 4), np.float32)
        vel = np.zeros((n, 4), np.float32)
        self._check(self.lib.cn_readback(self._ctx, pos.ctypes.data, prev.ctypes.data, vel.ctypes.data, n),
                    "cn_readback")
        return pos, prev, vel

    def meta(self, check: bool = False) -> abi.Meta:
        m = abi.Meta()
        st = self.lib.cn_validate_meta(self._ctx, C.byref(m))
        if check:
            self._check(st, "cn_validate_meta")
        elif st not in (abi.CN_OK, abi.CN_ERR_CAPACITY):
            self._check(st, "cn_validate_meta")
        return m

    def device_views(self) -> abi.DeviceViews:
        v = abi.DeviceViews()
        self._check(self.lib.cn_device_views_get(self._ctx, C.byref(v)), "cn_device_views_get")
        return v

    # -- zero-copy tensor views (the reference's shared_buffer_to_dlpack, python/bindings.cpp:282-340)
    def _acquire_view(self):
        self._nviews += 1

    def _release_view(self):
        self._nviews -= 1
        if self._nviews == 0 and self._close_pending:
            self._close_pending = False
            self.close()

    def dlpack_views(self) -> dict:
        """DLPack capsules of the resident particle state after the last step, particle-index order:
        ``positions_inv_mass`` and ``velocities``, float32 (count, 4).  The buffers are written on
        the solver's stream (``torch_stream``); they stay valid until ``reserve`` reallocates them
        (refused while views are alive) and keep the context alive past ``close``."""
        from . import dlpack
        v = self.device_views()
        if v.count == 0:
            raise RuntimeError("no resident scene")
        out = {}
        for name in ("positions_inv_mass", "velocities"):
            out[name] = dlpack.to_dlpack(getattr(v, name), (v.count, 4), "float32", dlpack.KDL_ROCM, self.device,
                                         owner=self, size_bytes=16 * v.count)
        return out

    def torch_views(self) -> dict:
        """``dlpack_views`` as torch tensors on the solver's device, without a copy."""
        import torch
        abi.require_single_hip_runtime("torch_views")
        return {k: torch.from_dlpack(c) for k, c in self.dlpack_views().items()}

    def set_scene_state(self, k: int, pos, vel):
        """Overwrite scene k's positions (xyz + inverse mass) and velocities, float32 (count, 4) in the
        scene's particle order: numpy arrays, or torch tensors on the solver's device (read in
        place after the current torch stream's work)."""
        first, count = C.c_uint32(), C.c_uint32()
        self._check(self.lib.cn_batch_scene_range(self._ctx, k, C.byref(first), C.byref(count)), "cn_batch_scene_range")
        if hasattr(pos, "data_ptr") and getattr(pos, "is_cuda", False):
            import torch
            abi.require_single_hip_runtime("set_scene_state")
            for t in (pos, vel):
                if t.dtype != torch.float32 or not t.is_contiguous() or tuple(t.shape) != (count.value, 4):
                    raise ValueError(f"state tensors must be contiguous float32 ({count.value}, 4)")
            lib_stream = self.torch_stream()
            lib_stream.wait_stream(torch.cuda.current_stream())
            self._check(self.lib.cn_scene_set_state(self._ctx, k, pos.data_ptr(), vel.data_ptr(),
                                                    abi.CN_STATE_DEVICE_POINTERS), "cn_scene_set_state")
            # later torch work (which may reuse the tensors' memory once they are freed) waits for
            # the library's read; no record_stream, which would outlive the context's stream
            torch.cuda.current_stream().wait_stream(lib_stream)
            return
        p = np.ascontiguousarray(pos, np.float32)
        v = np.ascontiguousarray(vel, np.float32)
        if p.shape != (count.value, 4) or v.shape != p.shape:
            raise ValueError(f"state arrays must be float32 ({count.value}, 4)")
        self._check(self.lib.cn_scene_set_state(self._ctx, k, p.ctypes.data, v.ctypes.data, 0), "cn_scene_set_state")

    def update_particles(self, first: int, pos, vel):
        """Dirty-run upload (cn_particles_update_range): overwrite the state of the uploaded
        particles [first, first + len(pos)) (cn_readback's order), leaving the rest resident."""
        p = np.ascontiguousarray(pos, np.float32).reshape(-1, 4)
        v = np.ascontiguousarray(vel, np.float32).reshape(-1, 4)
        if v.shape != p.shape:
            raise ValueError("positions and velocities must have the same shape")
        self._check(self.lib.cn_particles_update_range(self._ctx, first, len(p), p.ctypes.data if len(p) else None,
                                                       v.ctypes.data if len(v) else None, 0),
                    "cn_particles_update_range")

    def torch_stream(self):
        """The solver's HIP stream as a torch stream: order torch work after ``step`` with
        ``torch.cuda.current_stream().wait_stream(solver.torch_stream())``."""
        import torch
        abi.require_single_hip_runtime("torch_stream")
        return torch.cuda.ExternalStream(self.device_views().stream, device=torch.device("cuda", self.device))

    # -- per-pass entries
    def pass_keys(self, pos: np.ndarray, cell: float):
        n = pos.shape[0]
        pos = np.ascontiguousarray(pos, np.float32)
        keys = np.zeros(n, np.uint32)
        cells = np.zeros((n, 3), np.int32)
        self._check(self.lib.cn_pass_broadphase_keys(self._ctx, n, pos.ctypes.data, C.c_float(cell),
                                                     keys.ctypes.data, cells.ctypes.data), "keys")
        return keys, cells

    def pass_sort(self, keys: np.ndarray, idx: np.ndarray):
        k = np.ascontiguousarray(keys, np.uint32).copy()
        v = np.ascontiguousarray(idx, np.uint32).copy()
        self._check(self.lib.cn_pass_sort_pairs(self._ctx, len(k), k.ctypes.data, v.ctypes.data), "sort")
        return k, v

    def pass_sort_device_noalloc(self, keys_ptr: int, idx_ptr: int, count: int):
        """cn_pass_sort_pairs_noalloc: sorts count (key, index) pairs held in two device arrays in
        place, on the solver's stream, with no device allocation (the library's fixed scratch, at
        most CN_NOALLOC_MAX_PAIRS pairs).  Asynchronous: synchronize() before reading the result."""
        self._check(self.lib.cn_pass_sort_pairs_noalloc(self._ctx, count, keys_ptr, idx_ptr), "sort (no allocation)")

    def pass_cell_ranges(self, sorted_keys: np.ndarray, capacity: int):
        k = np.ascontiguousarray(sorted_keys, np.uint32)
        out = np.zeros((capacity, 4), np.uint32)
        self._check(self.lib.cn_pass_cell_ranges(self._ctx, len(k), k.ctypes.data, capacity, out.ctypes.data),
                    "cell ranges")
        return out

    def pass_substep_broadphase(self, params: abi.StepParams):
        n = self.n
        cap = params.max_fluid_neighbors
        keys = np.zeros(n, np.uint32)
        ids = np.zeros(n, np.uint32)
        counts = np.zeros(n, np.uint32)
        lists = np.zeros((n, max(cap, 1)), np.uint32)
        self._check(self.lib.cn_pass_substep_broadphase(self._ctx, C.byref(params), keys.ctypes.data,
                                                        ids.ctypes.data, counts.ctypes.data,
                                                        lists.ctypes.data), "substep broadphase")
        return keys, ids, counts, lists

    # -- profiling
    def profile(self, enable: bool):
        self._check(self.lib.cn_profile_enable(self._ctx, 1 if enable else 0), "profile")

    def profile_read(self) -> dict:
        s = abi.KernelStats()
        self._check(self.lib.cn_profile_read(self._ctx, C.byref(s)), "profile read")
        out = {}
        for k in range(s.used):
            name = bytes(s.name[k]).split(b"\0")[0].decode()
            out[name] = (int(s.launches[k]), float(s.total_ms[k]))
        return out
