"""ctypes front of the CPU oracle (liboracle: cn_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py, always as the checker or the CPU baseline, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcn_oracle.so")


class CnoScene(C.Structure):
    _fields_ = [("n", C.c_uint32), ("pos", C.c_void_p), ("prev", C.c_void_p), ("vel", C.c_void_p),
                ("radii", C.c_void_p), ("kinds", C.c_void_p), ("fluid_mat", C.c_void_p),
                ("part_mat", C.c_void_p), ("meta", C.c_void_p),
                ("n_fluid_mats", C.c_uint32), ("fluid_mats", C.c_void_p),
                ("n_part_mats", C.c_uint32), ("part_mats", C.c_void_p),
                ("n_colliders", C.c_uint32), ("colliders", C.c_void_p),
                ("n_edges", C.c_uint32), ("edges", C.c_void_p),
                ("n_bends", C.c_uint32), ("bends", C.c_void_p),
                ("n_dihedrals", C.c_uint32), ("dihedrals", C.c_void_p),
                ("n_tets", C.c_uint32), ("tets", C.c_void_p),
                ("edge_corr", C.c_void_p), ("bend_corr", C.c_void_p), ("tet_corr", C.c_void_p),
                ("render_aniso", C.c_void_p),
                ("adj_offsets", C.c_void_p), ("adj_counts", C.c_void_p), ("adj_indices", C.c_void_p),
                ("n_shape_colliders", C.c_uint32), ("shape_colliders", C.c_void_p)]


class CnoStats(C.Structure):
    _fields_ = [("max_required_neighbors", C.c_uint32), ("truncations", C.c_uint32),
                ("neighbor_total_first_iter", C.c_uint64), ("fluid_count", C.c_uint32),
                ("contact_overflows", C.c_uint32), ("max_abs_constraint_last", C.c_double),
                ("mean_constraint_last", C.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _P, _U = C.c_void_p, C.c_uint32
        sig = {
            "cno_default_params": (None, [_P]),
            "cno_fluid_material": (None, [_P, C.c_float, _P]),
            "cno_grid_cell_size": (C.c_float, [_U, _P, _P, _U, _P]),
            "cno_broadphase_entries": (None, [_U, _P, C.c_float, _P, _P]),
            "cno_radix_sort_1bit": (None, [_U, _P, _P]),
            "cno_next_pow2": (_U, [_U]),
            "cno_cell_ranges": (None, [_U, _P, _U, _P]),
            "cno_find_cell_range": (C.c_int, [_P, _U, _U, _P, _P]),
            "cno_static_collider_order": (None, [_U, _P, _P]),
            "cno_fluid_boundary_candidates": (_U, [_P, _P, _P, _P, _U]),
            "cno_fluid_neighbors": (None, [_P, C.c_float, _P, _P, _P, _P, _U, _P, _U, _U, _P, _P, _P]),
            "cno_predict": (None, [_P, C.c_float, _P]),
            "cno_fluid_density": (None, [_P, _P, _P, _U, _U, _P, _P, _P, _P, _P]),
            "cno_fluid_delta": (None, [_P, _P, C.c_float, _P, _U, _U, _P, _P, _P, _P, _P, _P]),
            "cno_fluid_apply": (None, [_P, _P, _U, _P, _P]),
            "cno_fluid_clamp": (None, [_P, _P, _P, _P, _P]),
            "cno_update_velocities": (None, [_P, C.c_float, _P]),
            "cno_project_boundary_velocities": (None, [_P, _P, C.c_float, _P, _P]),
            "cno_fluid_vorticity": (None, [_P, _U, _P, _P, _P]),
            "cno_fluid_vorticity_apply": (None, [_P, C.c_float, _U, _P, _P, _P]),
            "cno_fluid_render_anisotropy": (None, [_P, _U, _P, _P, _P]),
            "cno_step": (C.c_int, [_P, _P, _P]),
            "cno_threads": (C.c_int, []),
            "cno_set_threads": (None, [C.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


class OracleScene:
    """Owns contiguous copies of a cnhip.Scene's arrays in the oracle's layout."""

    def __init__(self, sc):
        self.pos = np.ascontiguousarray(sc.pos, np.float32).copy()
        self.prev = self.pos.copy()
        self.prev[:, 3] = 0.0
        self.vel = np.ascontiguousarray(sc.vel, np.float32).copy()
        self.radii = np.ascontiguousarray(sc.radii, np.float32)
        self.kinds = np.ascontiguousarray(sc.kinds, np.uint32)
        self.fluid_mat = np.ascontiguousarray(sc.fluid_mat, np.uint32)
        self.part_mat = np.ascontiguousarray(sc.part_mat, np.uint32)
        self.meta = sc.meta()
        self.fluid_mats = np.ascontiguousarray(sc.fluid_mats)
        self.part_mats = np.ascontiguousarray(sc.part_mats, np.float32)
        self.colliders = np.ascontiguousarray(sc.colliders)
        n = self.pos.shape[0]
        self.edges, self.bends, self.dihedrals, self.tets = sc.constraint_arrays()
        self.edge_corr = np.zeros((max(len(self.edges), 1), 8), np.float32)
        self.bend_corr = np.zeros((max(len(self.bends) + len(self.dihedrals), 1), 16), np.float32)
        self.tet_corr = np.zeros((max(len(self.tets), 1), 16), np.float32)

        def p(a):
            return a.ctypes.data if len(a) else None
        self.c = CnoScene(n, self.pos.ctypes.data, self.prev.ctypes.data, self.vel.ctypes.data,
                          self.radii.ctypes.data, self.kinds.ctypes.data, self.fluid_mat.ctypes.data,
                          self.part_mat.ctypes.data, self.meta.ctypes.data,
                          len(self.fluid_mats), self.fluid_mats.ctypes.data,
                          len(self.part_mats), self.part_mats.ctypes.data,
                          len(self.colliders), p(self.colliders),
                          len(self.edges), p(self.edges), len(self.bends), p(self.bends),
                          len(self.dihedrals), p(self.dihedrals), len(self.tets), p(self.tets),
                          self.edge_corr.ctypes.data, self.bend_corr.ctypes.data, self.tet_corr.ctypes.data)
        shapes = getattr(sc, "shape_colliders", None)
        if shapes is not None and len(shapes):
            self.shape_colliders = np.ascontiguousarray(shapes)
            self.c.n_shape_colliders = len(self.shape_colliders)
            self.c.shape_colliders = self.shape_colliders.ctypes.data
        adj = getattr(sc, "adjacency", None)
        if adj is not None:
            self.adj = tuple(np.ascontiguousarray(a, np.uint32) for a in adj)
            self.c.adj_offsets = self.adj[0].ctypes.data
            self.c.adj_counts = self.adj[1].ctypes.data
            self.c.adj_indices = self.adj[2].ctypes.data if len(self.adj[2]) else None

    @property
    def n(self):
        return self.pos.shape[0]

    def enable_render_anisotropy(self) -> np.ndarray:
        """Makes every step write the last substep's render anisotropy into the returned
        (3, n, 4) array: axis1, axis2, axis3 of every particle, w = scale."""
        self.render_aniso = np.zeros((3, self.n, 4), np.float32)
        self.c.render_aniso = self.render_aniso.ctypes.data
        return self.render_aniso


def step(osc: OracleScene, params) -> CnoStats:
    lib = load()
    st = CnoStats()
    rc = lib.cno_step(C.byref(osc.c), C.byref(params), C.byref(st))
    if rc != 0:
        raise RuntimeError("cno_step failed")
    return st


def fluid_material(radius: float, viscosity=0.01, cohesion=0.0, surface_tension=0.0, vorticity=0.0,
                   gravity_scale=1.0, cfl=1.0) -> np.ndarray:
    lib = load()
    desc = np.array([viscosity, cohesion, surface_tension, vorticity, gravity_scale, cfl], np.float32)
    out = np.zeros(12, np.float32)
    lib.cno_fluid_material(desc.ctypes.data, C.c_float(radius), out.ctypes.data)
    return out


def broadphase_entries(pos: np.ndarray, cell: float):
    lib = load()
    pos = np.ascontiguousarray(pos, np.float32)
    n = pos.shape[0]
    keys = np.zeros(n, np.uint32)
    cells = np.zeros((n, 3), np.int32)
    lib.cno_broadphase_entries(n, pos.ctypes.data, C.c_float(cell), keys.ctypes.data, cells.ctypes.data)
    return keys, cells


def radix_sort_1bit(keys: np.ndarray, idx: np.ndarray):
    lib = load()
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(idx, np.uint32).copy()
    lib.cno_radix_sort_1bit(len(k), k.ctypes.data, v.ctypes.data)
    return k, v


def next_pow2(v: int) -> int:
    return int(load().cno_next_pow2(v))


def cell_ranges(sorted_keys: np.ndarray, capacity: int) -> np.ndarray:
    lib = load()
    k = np.ascontiguousarray(sorted_keys, np.uint32)
    out = np.zeros((capacity, 4), np.uint32)
    lib.cno_cell_ranges(len(k), k.ctypes.data, capacity, out.ctypes.data)
    return out


def find_cell_range(ranges: np.ndarray, key: int):
    lib = load()
    r = np.ascontiguousarray(ranges, np.uint32)
    s = C.c_uint32()
    e = C.c_uint32()
    ok = lib.cno_find_cell_range(r.ctypes.data, r.shape[0], key, C.byref(s), C.byref(e))
    return (int(s.value), int(e.value)) if ok else None


def collider_contact(collider: np.ndarray, p):
    """(hit, signed distance, normal, aabb (6,)) of point p against one shape collider record."""
    lib = load()
    c = np.ascontiguousarray(collider).reshape(1)
    pt = np.ascontiguousarray(p, np.float32).reshape(3)
    sd = C.c_float()
    nrm = np.zeros(3, np.float32)
    aabb = np.zeros(6, np.float32)
    hit = lib.cno_collider_contact(C.c_void_p(c.ctypes.data), C.c_void_p(pt.ctypes.data), C.byref(sd),
                                   C.c_void_p(nrm.ctypes.data), C.c_void_p(aabb.ctypes.data))
    return bool(hit), float(sd.value), nrm, aabb


def rigid_contact(collider: np.ndarray, p, radius: float):
    """(hit, normal, penetration) of a solid particle at p with the given radius against one shape
    collider record (a static body of its own)."""
    lib = load()
    c = np.ascontiguousarray(collider).reshape(1)
    pt = np.ascontiguousarray(p, np.float32).reshape(3)
    pen = C.c_float()
    nrm = np.zeros(3, np.float32)
    hit = lib.cno_rigid_contact(C.c_void_p(c.ctypes.data), C.c_void_p(pt.ctypes.data), C.c_float(radius),
                                C.c_void_p(nrm.ctypes.data), C.byref(pen))
    return bool(hit), nrm, float(pen.value)


def static_collider_order(colliders: np.ndarray) -> np.ndarray:
    lib = load()
    c = np.ascontiguousarray(colliders)
    out = np.zeros(max(len(c), 1), np.uint32)
    lib.cno_static_collider_order(len(c), c.ctypes.data if len(c) else None, out.ctypes.data)
    return out[:len(c)]


def substep_neighbors(osc: OracleScene, params, cell: float):
    """Predict + entries + 1-bit sort + cell ranges + iteration-0 neighbor lists of one substep
    (does not modify osc)."""
    lib = load()
    tmp_pos = osc.pos.copy()
    tmp_vel = osc.vel.copy()
    tmp_prev = osc.prev.copy()
    c = CnoScene.from_buffer_copy(osc.c)
    c.pos, c.vel, c.prev = tmp_pos.ctypes.data, tmp_vel.ctypes.data, tmp_prev.ctypes.data
    dt = params.dt / max(params.substeps, 1)
    g = np.array(list(params.gravity)[:3], np.float32)
    lib.cno_predict(C.byref(c), C.c_float(dt), g.ctypes.data)
    keys, cells = broadphase_entries(tmp_pos, cell)
    idx = np.arange(osc.n, dtype=np.uint32)
    sk, si = radix_sort_1bit(keys, idx)
    cap_tbl = next_pow2(max(2 * osc.n, 1))
    tbl = cell_ranges(sk, cap_tbl)
    cap = params.max_fluid_neighbors
    counts = np.zeros(osc.n, np.uint32)
    lists = np.zeros((osc.n, max(cap, 1)), np.uint32)
    req = np.zeros(osc.n, np.uint32)
    dacc = np.zeros((osc.n, 4), np.float32)
    lib.cno_fluid_neighbors(C.byref(c), C.c_float(cell), cells.ctypes.data, sk.ctypes.data, si.ctypes.data,
                            tbl.ctypes.data, cap_tbl, dacc.ctypes.data, 0, cap, counts.ctypes.data,
                            lists.ctypes.data, req.ctypes.data)
    return dict(pos=tmp_pos, keys=keys, cells=cells, sorted_keys=sk, sorted_idx=si, table=tbl,
                counts=counts, lists=lists, required=req)


def fluid_density(osc: OracleScene, params, counts, lists, dacc=None, iteration=0,
                  bnd_ranges=None, bnd_pairs=None) -> np.ndarray:
    """physics_fluid_compute_density_lambda restated: (surface normal, C) per particle."""
    lib = load()
    n = osc.n
    cap = lists.shape[1]
    nc = np.zeros((n, 4), np.float32)
    dacc = np.zeros((n, 4), np.float32) if dacc is None else np.ascontiguousarray(dacc, np.float32)
    br = np.zeros((n, 2), np.uint32) if bnd_ranges is None else np.ascontiguousarray(bnd_ranges, np.uint32)
    bp = np.zeros(1, np.uint32) if bnd_pairs is None else np.ascontiguousarray(bnd_pairs, np.uint32)
    cfg = np.array(list(params.solver_config), np.float32)
    lists = np.ascontiguousarray(lists, np.uint32)
    counts = np.ascontiguousarray(counts, np.uint32)
    lib.cno_fluid_density(C.byref(osc.c), cfg.ctypes.data, dacc.ctypes.data, iteration, cap,
                          counts.ctypes.data, lists.ctypes.data, br.ctypes.data, bp.ctypes.data, nc.ctypes.data)
    return nc


def _boundary(n, bnd_ranges, bnd_pairs):
    br = np.zeros((n, 2), np.uint32) if bnd_ranges is None else np.ascontiguousarray(bnd_ranges, np.uint32)
    bp = np.zeros(1, np.uint32) if bnd_pairs is None or len(bnd_pairs) == 0 else \
        np.ascontiguousarray(bnd_pairs, np.uint32)
    return br, bp


def fluid_delta(osc: OracleScene, cfg, dt: float, counts, lists, nc, dacc=None, iteration=0,
                bnd_ranges=None, bnd_pairs=None) -> np.ndarray:
    """physics_fluid_compute_delta_positions restated: (delta, interaction weight) per particle.
    nc holds each particle's (surface normal, C) from the density pass; bnd_ranges (n, 2) the
    (offset, count) of each particle's static colliders in bnd_pairs."""
    lib = load()
    n = osc.n
    cfg = np.ascontiguousarray(cfg, np.float32)
    dacc = np.zeros((n, 4), np.float32) if dacc is None else np.ascontiguousarray(dacc, np.float32)
    br, bp = _boundary(n, bnd_ranges, bnd_pairs)
    lists = np.ascontiguousarray(lists, np.uint32)
    counts = np.ascontiguousarray(counts, np.uint32)
    nc = np.ascontiguousarray(nc, np.float32)
    out = np.zeros((n, 4), np.float32)
    lib.cno_fluid_delta(C.byref(osc.c), cfg.ctypes.data, C.c_float(dt), dacc.ctypes.data, iteration, lists.shape[1],
                        counts.ctypes.data, lists.ctypes.data, br.ctypes.data, bp.ctypes.data, nc.ctypes.data,
                        out.ctypes.data)
    return out


def fluid_apply(osc: OracleScene, cfg, iteration: int, din, dacc) -> np.ndarray:
    """physics_fluid_apply_delta_positions restated: the accumulated delta after one iteration."""
    lib = load()
    cfg = np.ascontiguousarray(cfg, np.float32)
    din = np.ascontiguousarray(din, np.float32)
    acc = np.ascontiguousarray(dacc, np.float32).copy()
    lib.cno_fluid_apply(C.byref(osc.c), cfg.ctypes.data, iteration, din.ctypes.data, acc.ctypes.data)
    return acc


def fluid_clamp(osc: OracleScene, dacc, bnd_ranges, bnd_pairs) -> np.ndarray:
    """physics_fluid_clamp_boundary restated: the accumulated delta pushed out of the colliders."""
    lib = load()
    din = np.ascontiguousarray(dacc, np.float32)
    br, bp = _boundary(osc.n, bnd_ranges, bnd_pairs)
    out = np.zeros_like(din)
    lib.cno_fluid_clamp(C.byref(osc.c), din.ctypes.data, br.ctypes.data, bp.ctypes.data, out.ctypes.data)
    return out


def update_velocities(osc: OracleScene, dt: float, dacc) -> np.ndarray:
    """physics_particle_update_velocities restated, in place on osc (pos, vel); returns the
    cleared delta buffer."""
    lib = load()
    acc = np.ascontiguousarray(dacc, np.float32).copy()
    lib.cno_update_velocities(C.byref(osc.c), C.c_float(dt), acc.ctypes.data)
    return acc


def project_boundary_velocities(osc: OracleScene, cfg, dt: float, bnd_ranges, bnd_pairs):
    """physics_fluid_project_boundary_velocities restated, in place on osc.vel."""
    lib = load()
    cfg = np.ascontiguousarray(cfg, np.float32)
    br, bp = _boundary(osc.n, bnd_ranges, bnd_pairs)
    lib.cno_project_boundary_velocities(C.byref(osc.c), cfg.ctypes.data, C.c_float(dt), br.ctypes.data,
                                        bp.ctypes.data)


def fluid_vorticity(osc: OracleScene, counts, lists) -> np.ndarray:
    """physics_fluid_compute_vorticity restated: (curl, |curl|) per particle."""
    lib = load()
    lists = np.ascontiguousarray(lists, np.uint32)
    counts = np.ascontiguousarray(counts, np.uint32)
    vort = np.zeros((osc.n, 4), np.float32)
    lib.cno_fluid_vorticity(C.byref(osc.c), lists.shape[1], counts.ctypes.data, lists.ctypes.data,
                            vort.ctypes.data)
    return vort


def fluid_vorticity_apply(osc: OracleScene, dt: float, counts, lists, vort):
    """physics_fluid_apply_vorticity_confinement restated: updates osc.vel in place."""
    lib = load()
    lists = np.ascontiguousarray(lists, np.uint32)
    counts = np.ascontiguousarray(counts, np.uint32)
    vort = np.ascontiguousarray(vort, np.float32)
    lib.cno_fluid_vorticity_apply(C.byref(osc.c), C.c_float(dt), lists.shape[1], counts.ctypes.data,
                                  lists.ctypes.data, vort.ctypes.data)


def fluid_render_anisotropy(osc: OracleScene, counts, lists) -> np.ndarray:
    """physics_fluid_build_render_anisotropy restated: (3, n, 4) = axis1..3 with their scales."""
    lib = load()
    lists = np.ascontiguousarray(lists, np.uint32)
    counts = np.ascontiguousarray(counts, np.uint32)
    out = np.zeros((3, osc.n, 4), np.float32)
    lib.cno_fluid_render_anisotropy(C.byref(osc.c), lists.shape[1], counts.ctypes.data, lists.ctypes.data,
                                    out.ctypes.data)
    return out


def threads() -> int:
    return int(load().cno_threads())


def set_threads(n: int):
    load().cno_set_threads(n)
