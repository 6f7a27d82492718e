"""Scene builders for the benchmark and test workloads (host authoring, particle-index order).

Fluid material and grid cell size follow the reference world: smoothing radius
h = max((2/0.6) r, 2r), rest density derived from the tight-pack kernel sum
(src/physics/physics_world.cpp:154-263), grid cell = max over particles of max(2r, h)
(:6829-6846).  Random numbers come from numpy's PCG64 with the stated seed.
"""
from __future__ import annotations

import numpy as np

from . import COLLIDER_DTYPE, Scene, fluid_material

CONTAINER_THICKNESS = 0.1


def container(lo, hi, thickness: float = CONTAINER_THICKNESS, closed: bool = True,
              material=(0.0, 0.0, 0.0, -1.0)) -> np.ndarray:
    """Static box walls around the interior [lo, hi] (floor, four sides and, if closed, a lid)."""
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    c = (lo + hi) * np.float32(0.5)
    he = (hi - lo) * np.float32(0.5)
    t = np.float32(thickness)
    walls = []
    for axis in range(3):
        for side in (-1, 1):
            if axis == 1 and side == 1 and not closed:
                continue
            center = c.copy()
            half = he + t  # walls overlap at the edges so the box has no gaps
            center[axis] = (lo[axis] - t * np.float32(0.5)) if side < 0 else (hi[axis] + t * np.float32(0.5))
            half = half.copy()
            half[axis] = t * np.float32(0.5)
            walls.append((center, half))
    out = np.zeros(len(walls), COLLIDER_DTYPE)
    for i, (center, half) in enumerate(walls):
        out[i]["center"] = center
        out[i]["half_extents"] = half
        out[i]["material"] = material
        out[i]["environment"] = 0
        out[i]["layer"] = 1
        out[i]["mask"] = 0xFFFFFFFF
        out[i]["enabled"] = 1
    return out


def _fluid_scene(pos_xyz: np.ndarray, radius: float, colliders: np.ndarray, name: str,
                 info: dict, particle_mass: float | None = None) -> Scene:
    n = pos_xyz.shape[0]
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = pos_xyz.astype(np.float32)
    mass = particle_mass if particle_mass is not None else 1.0
    pos[:, 3] = np.float32(1.0 / mass)
    fm = fluid_material(radius)
    return Scene(
        pos=pos, vel=np.zeros((n, 4), np.float32), radii=np.full(n, radius, np.float32),
        kinds=np.ones(n, np.uint32), fluid_mat=np.zeros(n, np.uint32), part_mat=np.zeros(n, np.uint32),
        env=np.zeros(n, np.uint32), phase=np.zeros(n, np.uint32), layer=np.ones(n, np.uint32),
        mask=np.full(n, 0xFFFFFFFF, np.uint32), fluid_mats=fm,
        part_mats=np.array([[0.0, 0.0, 0.0, -1.0]], np.float32), colliders=colliders, name=name, info=info)


def dam_break_64k(seed: int = 0, side: int = 40) -> Scene:
    """Plain PBF dam break: side^3 lattice at 0.05 m with +/-0.005 m seeded jitter, zero velocity,
    particle radius 0.03 m (h = 0.1 m = cell size), in the corner of a 4 x 2 x 2 m container.
    The 40^3 lattice spans 2 m per axis, so the container is 4 x 2 x 2 m rather than 2 x 1 x 1 m."""
    spacing = np.float32(0.05)
    g = (np.arange(side, dtype=np.float32) + np.float32(0.5)) * spacing
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float32)
    rng = np.random.Generator(np.random.PCG64(seed))
    p += rng.uniform(-0.005, 0.005, size=p.shape).astype(np.float32)
    cols = container((0.0, 0.0, 0.0), (4.0, 2.0, 2.0))
    return _fluid_scene(p, 0.03, cols, f"dam_break_{side ** 3}",
                        dict(spacing=0.05, jitter=0.005, seed=seed, box=(4.0, 2.0, 2.0)),
                        particle_mass=1000.0 * 0.05 ** 3)


def random_fill_radius(n: int, volume: float) -> float:
    """Radius whose tight-pack number density 1 / (0.75 (2r)^3) equals n / volume."""
    density = n / volume
    return float(0.5 * (0.75 * density) ** (-1.0 / 3.0))


def large_pbf(n: int = 4_194_304, seed: int = 2, box=(4.0, 2.0, 2.0)) -> Scene:
    """Large PBF: n particles i.i.d. uniform in a 4 x 2 x 2 m box at rest number density."""
    box = np.asarray(box, np.float64)
    r = random_fill_radius(n, float(np.prod(box)))
    rng = np.random.Generator(np.random.PCG64(seed))
    p = (rng.random((n, 3)) * box).astype(np.float32)
    cols = container((0.0, 0.0, 0.0), tuple(box))
    return _fluid_scene(p, r, cols, f"large_pbf_{n}", dict(seed=seed, box=tuple(box), radius=r))


SOFT_CONTACT = (0.0, 0.0, 0.0, -1.0)


def _solid_part(pos_xyz, radius, inv_mass, part_mat=0, env=0):
    n = len(pos_xyz)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = pos_xyz
    pos[:, 3] = np.float32(inv_mass)
    return dict(pos=pos, radii=np.full(n, radius, np.float32), kinds=np.zeros(n, np.uint32),
                fluid_mat=np.full(n, 0xFFFFFFFF, np.uint32), part_mat=np.full(n, part_mat, np.uint32),
                env=np.full(n, env, np.uint32))


def compose(parts, fluid_mats, colliders, name, info=None) -> Scene:
    """Concatenate pool parts (soft bodies, then cloths, then fluids -- the reference pool order of
    src/physics/physics_world.cpp:4930-5387) with their constraints shifted to pool indices."""
    from . import authoring as au
    n = sum(len(p["pos"]) for p in parts)
    cat = {k: np.concatenate([p[k] for p in parts]) for k in ("pos", "radii", "kinds", "fluid_mat", "part_mat", "env")}
    cons = {"edges": [], "bends": [], "dihedrals": [], "tets": []}
    adj = []
    off = 0
    for p in parts:
        if p.get("adj_pairs") is not None and len(p["adj_pairs"]):
            adj.append(np.asarray(p["adj_pairs"], np.int64) + off)
        for k in cons:
            if k in p and p[k] is not None and len(p[k]):
                a = p[k].copy()
                for f in {"edges": ("a", "b"), "bends": ("p0", "p1", "p2"),
                          "dihedrals": ("e0", "e1", "o0", "o1"), "tets": ("p",)}[k]:
                    a[f] += off
                cons[k].append(a)
        off += len(p["pos"])
    dts = {"edges": au.EDGE_DTYPE, "bends": au.BEND_DTYPE, "dihedrals": au.DIHEDRAL_DTYPE, "tets": au.TET_DTYPE}
    c = {k: (np.concatenate(v) if v else np.zeros(0, dts[k])) for k, v in cons.items()}
    # one phase group per part (packParticlePhase, include/physics/particle_phase.h:16-27): soft
    # bodies and cloths without self collision unless the part asks for it (self_collide), fluids
    # with the self-collide flag (physics_world.cpp:5083,5183,5381)
    phase = np.concatenate([np.full(len(p["pos"]), (k + 1) | (0x80000000 if (np.all(p["kinds"] == 1) or
                                                                            p.get("self_collide")) else 0), np.uint32)
                            for k, p in enumerate(parts)])
    return Scene(pos=cat["pos"], vel=np.zeros((n, 4), np.float32), radii=cat["radii"], kinds=cat["kinds"],
                 fluid_mat=cat["fluid_mat"], part_mat=cat["part_mat"], env=cat["env"],
                 phase=phase, layer=np.ones(n, np.uint32), mask=np.full(n, 0xFFFFFFFF, np.uint32),
                 fluid_mats=fluid_mats, part_mats=np.array([SOFT_CONTACT], np.float32), colliders=colliders,
                 name=name, info=info or {}, edges=c["edges"], bends=c["bends"], dihedrals=c["dihedrals"],
                 tets=c["tets"], adjacency=au.adjacency(n, np.concatenate(adj)) if adj else None)


def soft_block_part(res=8, spacing=0.05, radius=None, translation=(1.0, 0.6, 1.0), mass=0.1, compliance=1e-6):
    from . import authoring as au
    size = (res - 1) * spacing
    pos, tets = au.tet_grid((size, size, size), resolution=(res, res, res), translation=translation)
    edges, tc = au.tet_constraints(pos, tets, 0, compliance, compliance)
    part = _solid_part(pos, radius if radius is not None else 0.5 * spacing, 1.0 / mass)
    part.update(edges=edges, tets=tc, adj_pairs=au.tet_adjacency_pairs(edges))
    return part


def cloth_part(n=32, spacing=0.02, radius=None, translation=(0.5, 1.5, 0.5), mass=0.01, compliance=1e-6):
    from . import authoring as au
    pos, tris = au.cloth_grid(n, n, spacing, translation)
    edges, dih = au.cloth_constraints(pos, tris, 0, compliance, compliance)
    part = _solid_part(pos, radius if radius is not None else 0.5 * spacing, 1.0 / mass)
    part.update(edges=edges, dihedrals=dih, adj_pairs=au.cloth_adjacency_pairs(edges, dih))
    return part


def fluid_part(side=16, spacing=0.05, radius=0.03, origin=(0.0, 0.0, 0.0), seed=0, jitter=0.0, fluid_mat=0):
    g = (np.arange(side, dtype=np.float32) + np.float32(0.5)) * np.float32(spacing)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float32) + np.asarray(origin, np.float32)
    if jitter:
        rng = np.random.Generator(np.random.PCG64(seed))
        p += rng.uniform(-jitter, jitter, size=p.shape).astype(np.float32)
    n = len(p)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = p
    pos[:, 3] = np.float32(1.0 / (1000.0 * spacing ** 3))
    return dict(pos=pos, radii=np.full(n, radius, np.float32), kinds=np.ones(n, np.uint32),
                fluid_mat=np.full(n, fluid_mat, np.uint32), part_mat=np.zeros(n, np.uint32),
                env=np.zeros(n, np.uint32))


def mixed(fluid_side=64, tet_res=32, cloth_n=256, seed=1) -> Scene:
    """Mixed single scene: a fluid block (fluid_side^3 particles, r = 0.01 m), a tet soft block from a
    tet_res^3 vertex grid (5 tets per cube) and a cloth_n^2 cloth, compliance 1e-6, in a 4 x 2 x 2 m
    container (pool order: soft body, cloth, fluid)."""
    r = 0.01
    soft = soft_block_part(tet_res, 0.02, r, translation=(3.2, 0.4, 1.0))
    cloth = cloth_part(cloth_n, 1.9 / max(cloth_n - 1, 1), r * 0.5, translation=(0.05, 1.6, 0.05))
    fluid = fluid_part(fluid_side, 0.02, r, origin=(0.05, 0.02, 0.05), seed=seed, jitter=0.002)
    fm = fluid_material(r)
    cols = container((0.0, 0.0, 0.0), (4.0, 2.0, 2.0))
    return compose([soft, cloth, fluid], fm, cols, f"mixed_{fluid_side}_{tet_res}_{cloth_n}",
                   dict(fluid=fluid_side ** 3, soft=tet_res ** 3, cloth=cloth_n ** 2))


def contact_stack(block_res=6, cloth_n=24, fluid_side=6) -> Scene:
    """Particle-particle contact workload: a tet soft block (spacing 0.05 m, r = 0.025 m) lying on a
    cloth (spacing 0.03 m, r = 0.015 m) with a fluid block (spacing 0.05 m, r = 0.03 m) on top, every
    layer overlapping the next by 0.01 m at the start, in a 4 x 2 x 2 m container."""
    cloth = cloth_part(cloth_n, 0.03, 0.015, translation=(0.7, 0.30, 0.7))
    half = 0.5 * (block_res - 1) * 0.05
    c = 0.7 + 0.5 * (cloth_n - 1) * 0.03  # block centred over the cloth (the tet grid is centred)
    soft = soft_block_part(block_res, 0.05, 0.025, translation=(c, 0.30 + 0.03 + half, c))
    top = 0.30 + 0.03 + 2 * half
    # fluid columns over the block's vertex columns (fluid_part's lattice starts half a spacing in)
    fluid = fluid_part(fluid_side, 0.05, 0.03, origin=(c - half - 0.025, top + 0.045 - 0.025, c - half - 0.025))
    cols = container((0.0, 0.0, 0.0), (4.0, 2.0, 2.0))
    return compose([soft, cloth, fluid], fluid_material(0.03), cols,
                   f"contact_stack_{block_res}_{cloth_n}_{fluid_side}",
                   dict(soft=block_res ** 3, cloth=cloth_n ** 2, fluid=fluid_side ** 3))


def obstacles(material=(0.0, 0.0, 0.0, -1.0)) -> np.ndarray:
    """Three static shape colliders on the floor of the 4 x 2 x 2 m container: a sphere (r 0.12 m),
    a capsule lying along z (r 0.06 m, half height 0.2 m) and a box turned 30 degrees about y."""
    from . import shape_collider
    s = np.float32(np.sin(np.pi / 4)), np.float32(np.sin(np.pi / 12))
    c = np.float32(np.cos(np.pi / 4)), np.float32(np.cos(np.pi / 12))
    return np.concatenate([
        shape_collider(0, (0.25, 0.15, 0.25), params=(0.12,), material=material),
        shape_collider(2, (0.5, 0.1, 0.35), orientation=(s[0], 0.0, 0.0, c[0]), params=(0.06, 0.2),
                       material=material),
        shape_collider(1, (0.3, 0.08, 0.5), orientation=(0.0, s[1], 0.0, c[1]), params=(0.1, 0.05, 0.1),
                       material=material),
    ])


def shapes_scene(side: int = 10, seed: int = 0) -> Scene:
    """A jittered fluid block (spacing 0.05 m, r = 0.03 m) released above the obstacles() in the
    4 x 2 x 2 m container."""
    part = fluid_part(side, 0.05, 0.03, origin=(0.05, 0.35, 0.05), seed=seed, jitter=0.005)
    cols = container((0.0, 0.0, 0.0), (4.0, 2.0, 2.0))
    sc = compose([part], fluid_material(0.03), cols, f"shapes_{side ** 3}", dict(side=side))
    sc.shape_colliders = obstacles()
    return sc


def pegboard_colliders(nx: int = 10, nz: int = 6, y: float = 0.25, spacing: float = 0.12,
                       origin=(0.2, 0.2)) -> np.ndarray:
    """A grid of nx x nz static pegs at height y, cycling sphere (r 0.03 m), capsule lying along x
    (r 0.02 m, half height 0.03 m) and box turned about y (half extents 0.03 x 0.02 x 0.025 m)."""
    from . import shape_collider
    out = []
    for i in range(nx):
        for j in range(nz):
            c = (origin[0] + i * spacing, y, origin[1] + j * spacing)
            kind = (i + j) % 3
            if kind == 0:
                out.append(shape_collider(0, c, params=(0.03,)))
            elif kind == 1:
                h = np.float32(np.sqrt(0.5))
                out.append(shape_collider(2, c, orientation=(0.0, 0.0, h, h), params=(0.02, 0.03)))
            else:
                a = 0.3 + 0.1 * i
                out.append(shape_collider(1, c, orientation=(0.0, np.sin(a / 2), 0.0, np.cos(a / 2)),
                                          params=(0.03, 0.02, 0.025)))
    return np.concatenate(out)


def pegboard(fluid_side: int = 10, cloth_n: int = 12, nx: int = 10, nz: int = 6) -> Scene:
    """A fluid block and a cloth dropped onto a board of nx x nz pegs (more than 32 static colliders
    with the container's six walls) in the 4 x 2 x 2 m container."""
    spacing = 0.12
    span = (nx - 1) * spacing
    fluid = fluid_part(fluid_side, 0.04, 0.02, origin=(0.2, 0.4, 0.2), seed=3, jitter=0.004)
    cloth = cloth_part(cloth_n, 0.04, 0.02, translation=(0.2 + 0.5 * span, 0.9, 0.2 + 0.3))
    cols = container((0.0, 0.0, 0.0), (4.0, 2.0, 2.0))
    sc = compose([cloth, fluid], fluid_material(0.02), cols, f"pegboard_{nx}x{nz}",
                 dict(fluid=fluid_side ** 3, cloth=cloth_n ** 2, pegs=nx * nz))
    sc.shape_colliders = pegboard_colliders(nx, nz, spacing=spacing)
    return sc


def folded_cloth(n: int = 16, spacing: float = 0.03, gap: float = 0.75, self_collide: bool = True) -> Scene:
    """A self-colliding cloth folded onto itself: an n x n grid (spacing 0.03 m, r = spacing / 2)
    authored flat, then the half beyond the middle column mirrored over it, gap x 2r above, so
    every upper-layer particle overlaps the one below it.  Across the crease the mirrored columns
    are hinge opposites, so constraint adjacency keeps those pairs apart; the others collide."""
    cloth = cloth_part(n, spacing, 0.5 * spacing, translation=(1.0, 0.5, 1.0))
    cloth["self_collide"] = self_collide
    p = cloth["pos"]
    mid = np.float32(1.0 + (n // 2) * spacing - 0.5 * (n - 1) * spacing)  # x of the middle column
    xs = np.unique(p[:, 0])
    mid = xs[n // 2] if len(xs) == n else mid
    over = p[:, 0] > mid
    p[over, 0] = np.float32(2.0) * mid - p[over, 0]
    p[over, 1] = p[over, 1] + np.float32(gap * spacing)
    cols = container((0.0, 0.0, 0.0), (4.0, 2.0, 2.0))
    return compose([cloth], fluid_material(0.03), cols, f"folded_cloth_{n}", dict(n=n, self_collide=self_collide))


def scaled_box_for(n: int, n_ref: int = 4_194_304, box_ref=(4.0, 2.0, 2.0)):
    """Box with the same aspect and number density as box_ref holds for n_ref particles."""
    s = (n / n_ref) ** (1.0 / 3.0)
    return tuple(float(b * s) for b in box_ref)


def batch_sizes(count: int = 512, seed: int = 3, lo: int = 2000, hi: int = 64000):
    """Particle counts log-uniform in [lo, hi] and scene kinds (0 PBF 40 %, 1 tet soft body 30 %,
    2 cloth 30 %) of a batch of independent scenes, from one seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)
    kind = rng.choice(3, size=count, p=[0.4, 0.3, 0.3])
    return n, kind


def batch_scene(n: int, kind: int, seed: int = 0) -> Scene:
    """One scene of the batched workload with about n particles in its own 4 x 2 x 2 m container:
    a jittered PBF block (spacing 0.05 m, r = 0.03 m, the dam-break geometry), a tet soft block
    (spacing 0.04 m) 0.05 m above the floor, or a cloth grid spanning 1.9 m, compliance 1e-6."""
    cols = container((0.0, 0.0, 0.0), (4.0, 2.0, 2.0))
    fm = fluid_material(0.03)
    if kind == 0:
        side = max(2, int(round(n ** (1.0 / 3.0))))
        part = fluid_part(side, 0.05, 0.03, origin=(0.05, 0.02, 0.05), seed=seed, jitter=0.005)
    elif kind == 1:
        res = max(2, int(round(n ** (1.0 / 3.0))))
        half = 0.5 * (res - 1) * 0.04  # the tet grid is centred on its translation
        part = soft_block_part(res, 0.04, 0.02, translation=(2.0, half + 0.05, 1.0))
    else:
        m = max(2, int(round(n ** 0.5)))
        sp = 1.9 / (m - 1)
        part = cloth_part(m, sp, 0.5 * sp, translation=(0.05, 1.2, 0.05))
    return compose([part], fm, cols, f"batch_{('pbf', 'tet', 'cloth')[kind]}_{n}", dict(n=int(n), kind=int(kind)))


def _batch_scene_args(args):
    return batch_scene(*args)


def batched(count: int = 512, seed: int = 3, lo: int = 2000, hi: int = 64000, workers: int = 1):
    """The batched workload: count independent scenes sized and typed by batch_sizes().  workers > 1
    builds them in that many fresh (spawned) processes; the scenes are the same."""
    n, kind = batch_sizes(count, seed, lo, hi)
    args = [(int(n[k]), int(kind[k]), k) for k in range(count)]
    if workers <= 1:
        return [batch_scene(*a) for a in args]
    import multiprocessing as mp
    with mp.get_context("spawn").Pool(workers) as pool:
        return pool.map(_batch_scene_args, args, chunksize=max(1, count // (4 * workers)))
