"""Triangle meshes for the mesh collider (pbf_set_mesh): the scene geometry
the reference ray-casts particles against (src/particles.cpp:51-132).  Each
mesh is (pos[nv, 3], nrm[nv, 3], tri[nt, 3]); vertex normals face the fluid,
as the reference's scene meshes do for the slide test dot(d, n) > -1."""
from __future__ import annotations

import numpy as np


def box(lo, hi, skip=()):
    """The inside of an axis-aligned box as 12 triangles with inward normals
    (the Cornell box walls).  skip: faces (2*axis + side) left out."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    pos, nrm, tri = [], [], []
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        for side in (0, 1):
            if 2 * a + side in skip:
                continue
            plane = hi[a] if side else lo[a]
            n = np.zeros(3)
            n[a] = -1.0 if side else 1.0
            base = len(pos)
            for u, v in ((0, 0), (1, 0), (1, 1), (0, 1)):
                q = np.zeros(3)
                q[a] = plane
                q[b] = hi[b] if u else lo[b]
                q[c] = hi[c] if v else lo[c]
                pos.append(q)
                nrm.append(n)
            tri += [[base, base + 1, base + 2], [base, base + 2, base + 3]]
    return np.array(pos), np.array(nrm), np.array(tri, np.int32)


def icosphere(center, radius, subdiv=2):
    """A sphere of 20 * 4^subdiv triangles with outward vertex normals (an
    obstacle the fluid hits from outside)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
         [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
         [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
         [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        mid = {}

        def midpoint(i, j):
            key = (min(i, j), max(i, j))
            if key not in mid:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                mid[key] = len(verts) - 1
            return mid[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    unit = np.array(verts)
    return np.asarray(center, float) + radius * unit, unit.copy(), np.array(f, np.int32)


def default_skin(pos, box_lo, box_hi) -> float:
    """A float32 collision skin for pbf_set_mesh: 2^-17 of the largest
    coordinate of the mesh and the box (128x the float rounding there)."""
    scale = max(float(np.abs(pos).max()), float(np.abs(box_lo).max()), float(np.abs(box_hi).max()))
    return float(np.float32(scale * 2.0 ** -17))


def merge(*meshes):
    """One mesh from several (indices shifted)."""
    pos, nrm, tri, off = [], [], [], 0
    for p, n, t in meshes:
        pos.append(p)
        nrm.append(n)
        tri.append(np.asarray(t) + off)
        off += len(p)
    return np.concatenate(pos), np.concatenate(nrm), np.concatenate(tri).astype(np.int32)
