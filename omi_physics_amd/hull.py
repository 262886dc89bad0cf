"""Convex-hull cooking for ``convex`` colliders (host, set-up time).

The GPU's convex support function returns the first hull vertex of maximal
projection, so the *order* of the hull vertices is part of the contact
result.  This module therefore rebuilds hulls with the reference's
incremental algorithm step for step (``hull.py:16-102``): unique points,
seed tetrahedron from the farthest point / largest area / largest height,
then each point in turn replaces the faces it sees with fans over their
horizon; the used vertices are returned in index order.  The parity test
checks the vertex arrays against the reference's on 80 random clouds.
"""
from typing import Dict, List, Optional, Tuple

import numpy as np

EPS = 1e-9
Face = Tuple[int, int, int]


def _seed(pts: np.ndarray) -> Optional[List[Face]]:
    """Outward faces of the seed tetrahedron, or None for degenerate input."""
    p0 = pts[0]
    i1 = int(np.argmax(np.linalg.norm(pts - p0, axis=1)))
    if np.linalg.norm(pts[i1] - p0) < EPS:
        return None
    edge = pts[i1] - p0
    area = np.linalg.norm(np.cross(pts - p0, edge), axis=1)
    i2 = int(np.argmax(area))
    if area[i2] < EPS:
        return None
    nrm = np.cross(pts[i1] - p0, pts[i2] - p0)
    height = np.abs((pts - p0) @ nrm)
    i3 = int(np.argmax(height))
    if height[i3] < EPS:
        return None
    a, b, c, d = 0, i1, i2, i3
    if np.dot(np.cross(pts[b] - pts[a], pts[c] - pts[a]), pts[d] - pts[a]) > 0:
        b, c = c, b
    return [(a, b, c), (a, c, d), (a, d, b), (b, d, c)]


def _plane(pts: np.ndarray, f: Face) -> Tuple[np.ndarray, np.ndarray]:
    a = pts[f[0]]
    n = np.cross(pts[f[1]] - a, pts[f[2]] - a)
    ln = np.linalg.norm(n)
    return (n / ln if ln > EPS else n), a


def _horizon(visible: List[Face]) -> List[Tuple[int, int]]:
    """Directed boundary edges of the visible patch, in face order."""
    count: Dict[Tuple[int, int], int] = {}
    for f in visible:
        for e in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            k = (min(e), max(e))
            count[k] = count.get(k, 0) + 1
    return [e for f in visible for e in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0]))
            if count[(min(e), max(e))] == 1]


def convex_hull(points) -> Tuple[np.ndarray, np.ndarray]:
    """``(vertices, triangles)`` of the hull; degenerate input returns the unique points.

    Each face's unit normal and reference vertex are computed once, when the face
    is created, and every point tests all current faces in one array operation.
    """
    pts = np.unique(np.asarray(points, dtype='d'), axis=0)
    if len(pts) < 4:
        return pts, np.zeros((0, 3), dtype='i')
    faces = _seed(pts)
    if faces is None:
        return pts, np.zeros((0, 3), dtype='i')
    planes = [_plane(pts, f) for f in faces]
    nrm = np.array([pl[0] for pl in planes])
    ref = np.array([pl[1] for pl in planes])
    for p in range(len(pts)):
        sees = ((pts[p] - ref) * nrm).sum(axis=1) > EPS
        if not sees.any():
            continue
        visible = [f for f, v in zip(faces, sees) if v]
        new = [(i, j, p) for i, j in _horizon(visible)]
        keep = ~sees
        faces = [f for f, v in zip(faces, keep) if v] + new
        planes = [_plane(pts, f) for f in new]
        nrm = np.concatenate([nrm[keep], np.array([pl[0] for pl in planes]).reshape(-1, 3)])
        ref = np.concatenate([ref[keep], np.array([pl[1] for pl in planes]).reshape(-1, 3)])
    used = np.unique(np.asarray(faces).ravel())
    remap = {int(o): k for k, o in enumerate(used)}
    tris = np.array([[remap[i] for i in f] for f in faces], dtype='i')
    return pts[used], tris


def cook_convex(points) -> np.ndarray:
    """Hull vertices for a ``convex`` shape, or the raw points if degenerate (cookery.py:63-68)."""
    verts, _ = convex_hull(points)
    return verts if len(verts) >= 4 else np.asarray(points, dtype='d')
