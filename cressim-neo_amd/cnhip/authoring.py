"""Host authoring of soft bodies and cloth (inputs of the particle path), float32 like the reference.

Tet grid: buildRegularGridSource + the edge/volume derivation of src/physics/soft_body_authoring.cpp
:81-177, :319-432 and the pool build of src/physics/physics_world.cpp:5090-5141 (edges in
first-seen order over the tets' (0,1),(0,2),(0,3),(1,2),(1,3),(2,3) pairs, rest length
sqrt(dot(d, d)), rest volume |dot(cross(p1-p0, p2-p0), p3-p0)| / 6).
Cloth: cookClothTopology of src/physics/cloth_authoring.cpp:55-191 (structural edges in
ascending (lo, hi) order, a dihedral hinge for every edge with two opposite vertices, opposites
sorted, rest angle from signedDihedral).
"""
from __future__ import annotations

import math

import numpy as np

F = np.float32

EDGE_DTYPE = np.dtype([("a", np.uint32), ("b", np.uint32), ("rest_length", np.float32),
                       ("compliance", np.float32), ("enabled", np.uint32), ("reserved", np.uint32, 3)])
BEND_DTYPE = np.dtype([("p0", np.uint32), ("p1", np.uint32), ("p2", np.uint32), ("rest_angle", np.float32),
                       ("compliance", np.float32), ("reserved", np.uint32, 3)])
DIHEDRAL_DTYPE = np.dtype([("e0", np.uint32), ("e1", np.uint32), ("o0", np.uint32), ("o1", np.uint32),
                           ("rest_angle", np.float32), ("compliance", np.float32), ("reserved", np.uint32, 2)])
TET_DTYPE = np.dtype([("p", np.uint32, 4), ("rest_volume", np.float32), ("compliance", np.float32),
                      ("reserved", np.uint32, 2)])
assert EDGE_DTYPE.itemsize == 32 and BEND_DTYPE.itemsize == 32 and DIHEDRAL_DTYPE.itemsize == 32
assert TET_DTYPE.itemsize == 32


def _dot(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)


def tet_grid(size=(1.0, 1.0, 1.0), spacing=0.5, translation=(0.0, 0.0, 0.0), resolution=None):
    """Regular-grid tet mesh: positions (float32, particle order x-major), tets (m, 4) uint32."""
    size = np.maximum(np.asarray(size, np.float32), F(1e-4))
    if resolution is None:
        res = [max(2, int(math.ceil(float(F(e) / F(spacing)))) + 1) for e in size]
    else:
        res = list(resolution)
    rx, ry, rz = res
    tx = (np.arange(rx, dtype=np.float32) / F(rx - 1)) if rx > 1 else np.zeros(1, np.float32)
    ty = (np.arange(ry, dtype=np.float32) / F(ry - 1)) if ry > 1 else np.zeros(1, np.float32)
    tz = (np.arange(rz, dtype=np.float32) / F(rz - 1)) if rz > 1 else np.zeros(1, np.float32)
    X, Y, Z = np.meshgrid((tx - F(0.5)) * size[0], (ty - F(0.5)) * size[1], (tz - F(0.5)) * size[2],
                          indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32) + np.asarray(translation, np.float32)

    def flat(x, y, z):
        return (x * ry + y) * rz + z

    # cubes in x, y, z order (z fastest), five tets each, the split alternating with (x + y + z) parity
    x, y, z = (v.ravel().astype(np.int64) for v in
               np.meshgrid(np.arange(rx - 1), np.arange(ry - 1), np.arange(rz - 1), indexing="ij"))
    p0, p1 = flat(x, y, z), flat(x, y, z + 1)
    p3, p2 = flat(x + 1, y, z), flat(x + 1, y, z + 1)
    p7, p6 = flat(x + 1, y + 1, z), flat(x + 1, y + 1, z + 1)
    p4, p5 = flat(x, y + 1, z), flat(x, y + 1, z + 1)
    odd = np.stack([np.stack(t, -1) for t in ((p2, p1, p6, p3), (p6, p3, p4, p7), (p4, p1, p6, p5),
                                               (p3, p1, p4, p0), (p6, p1, p4, p3))], 1)
    even = np.stack([np.stack(t, -1) for t in ((p0, p2, p5, p1), (p7, p2, p0, p3), (p5, p2, p7, p6),
                                                (p7, p0, p5, p4), (p0, p2, p7, p5))], 1)
    tets = np.where(((x + y + z) & 1).astype(bool)[:, None, None], odd, even)
    return pos, tets.reshape(-1, 4).astype(np.uint32)


def tet_constraints(pos, tets, offset=0, edge_compliance=0.0, volume_compliance=0.0):
    """Edges (first-seen order) and tet volume constraints with float32 rest values."""
    pairs = np.array([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)])
    e = tets[:, pairs].reshape(-1, 2).astype(np.int64)
    lo, hi = np.minimum(e[:, 0], e[:, 1]), np.maximum(e[:, 0], e[:, 1])
    key = (lo << 32) | hi
    _, first = np.unique(key, return_index=True)
    first.sort()
    lo, hi = lo[first], hi[first]
    d = pos[hi] - pos[lo]
    edges = np.zeros(len(lo), EDGE_DTYPE)
    edges["a"] = lo + offset
    edges["b"] = hi + offset
    edges["rest_length"] = np.sqrt(_dot(d, d).astype(np.float32))
    edges["compliance"] = edge_compliance
    edges["enabled"] = 1
    p0, p1, p2, p3 = (pos[tets[:, k]] for k in range(4))
    vol = np.abs(_dot(_cross(p1 - p0, p2 - p0), p3 - p0)) / F(6.0)
    tc = np.zeros(len(tets), TET_DTYPE)
    tc["p"] = tets + offset
    tc["rest_volume"] = vol.astype(np.float32)
    tc["compliance"] = volume_compliance
    return edges, tc


def cloth_grid(nx, ny, spacing, translation=(0.0, 0.0, 0.0)):
    """Flat cloth in the xz plane: positions (nx * ny, 3) and triangles (2 per quad)."""
    g = np.arange(nx, dtype=np.float32) * F(spacing)
    h = np.arange(ny, dtype=np.float32) * F(spacing)
    Z, X = np.meshgrid(h, g, indexing="ij")
    pos = np.stack([X.ravel(), np.zeros(nx * ny, np.float32), Z.ravel()], 1) + np.asarray(translation, np.float32)
    i = np.arange(ny - 1)[:, None] * nx + np.arange(nx - 1)[None, :]
    a, b, c, d = i, i + 1, i + nx, i + nx + 1
    tris = np.stack([np.stack([a, b, d], -1), np.stack([a, d, c], -1)], 2).reshape(-1, 3).astype(np.uint32)
    return pos.astype(np.float32), tris


def _signed_dihedral(x0, x1, x2, x3):
    def normalize(v):
        return v / np.float32(math.sqrt(float(_dot(v, v))))
    edge = normalize(x1 - x0)
    n0 = normalize(_cross(edge, x2 - x0))
    n1 = normalize(_cross(x3 - x0, edge))
    return F(math.atan2(float(_dot(edge, _cross(n0, n1))), float(_dot(n0, n1))))


def cloth_constraints(pos, tris, offset=0, structural_compliance=0.0, bend_compliance=0.0):
    recs = np.concatenate([tris[:, [0, 1, 2]], tris[:, [1, 2, 0]], tris[:, [2, 0, 1]]], 0).astype(np.int64)
    order = np.arange(len(tris))
    tri_of = np.concatenate([order, order, order])
    rec_of = np.repeat(np.arange(3), len(tris))
    lo, hi = np.minimum(recs[:, 0], recs[:, 1]), np.maximum(recs[:, 0], recs[:, 1])
    key = (lo << 32) | hi
    # opposites are appended in triangle order, record order inside a triangle
    seq = tri_of * 3 + rec_of
    o = np.lexsort((seq, key))
    key, lo, hi, opp = key[o], lo[o], hi[o], recs[o, 2]
    uk, start, cnt = np.unique(key, return_index=True, return_counts=True)
    elo, ehi = lo[start], hi[start]
    d = pos[ehi] - pos[elo]
    edges = np.zeros(len(uk), EDGE_DTYPE)
    edges["a"] = elo + offset
    edges["b"] = ehi + offset
    edges["rest_length"] = np.sqrt(_dot(d, d).astype(np.float32))
    edges["compliance"] = structural_compliance
    edges["enabled"] = 1
    two = np.nonzero(cnt == 2)[0]
    c0, c1 = opp[start[two]], opp[start[two] + 1]
    oc, od = np.minimum(c0, c1), np.maximum(c0, c1)
    dih = np.zeros(len(two), DIHEDRAL_DTYPE)
    dih["e0"] = elo[two] + offset
    dih["e1"] = ehi[two] + offset
    dih["o0"] = oc + offset
    dih["o1"] = od + offset
    dih["compliance"] = bend_compliance
    # flat cloth: every rest angle is atan2(0, 1) = 0; general meshes evaluate signedDihedral
    p = pos
    flat = np.allclose(p[:, 1], p[0, 1])
    if not flat:
        dih["rest_angle"] = [_signed_dihedral(p[a], p[b], p[c], p[e])
                             for a, b, c, e in zip(elo[two], ehi[two], oc, od)]
    return edges, dih


def adjacency(n: int, pairs) -> tuple:
    """Constraint adjacency as the reference builds it: every pair in both directions, each
    particle's list sorted and deduplicated (soft_body_authoring.cpp:392-408 for tet edges,
    cloth_authoring.cpp:139-188 for structural edges and dihedral-hinge opposites), laid out as
    the pool's offsets / counts / indices (physics_world.cpp:5774-5787).  pairs: (m, 2)."""
    pr = np.asarray(pairs, np.int64).reshape(-1, 2)
    a = np.concatenate([pr[:, 0], pr[:, 1]])
    b = np.concatenate([pr[:, 1], pr[:, 0]])
    key = np.unique(a * np.int64(max(n, 1)) + b)
    a, b = key // max(n, 1), key % max(n, 1)
    counts = np.bincount(a, minlength=n).astype(np.uint32)
    offsets = np.zeros(n, np.uint32)
    if n > 1:
        offsets[1:] = np.cumsum(counts[:-1], dtype=np.uint64).astype(np.uint32)
    return offsets, counts, b.astype(np.uint32)


def tet_adjacency_pairs(edges) -> np.ndarray:
    """The pairs of a tet soft body's adjacency: its tet edges."""
    return np.stack([edges["a"], edges["b"]], 1).astype(np.int64)


def cloth_adjacency_pairs(edges, dihedrals) -> np.ndarray:
    """The pairs of a cloth's adjacency: structural edges and the two opposite vertices of every
    bending hinge."""
    return np.concatenate([np.stack([edges["a"], edges["b"]], 1), np.stack([dihedrals["o0"], dihedrals["o1"]], 1)]
                          ).astype(np.int64)
