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


def scaled_box_for(n: int, n_ref: int = 4_194_304, box_ref=(4.0, 2.0, 2.0)):
    """Box with the same aspect and number density as box_ref holds for n_ref particles."""
    s = (n / n_ref) ** (1.0 / 3.0)
    return tuple(float(b * s) for b in box_ref)
