"""Capsule against a batch of triangles on the GPU (the character controller's query).

:func:`capsule_triangles` and :func:`capsule_triangle_batch` answer
``collide.capsule_triangles`` / ``collide.capsule_triangle_batch``
(reference ``collide.py:385-462``) through ``omi_capsule_triangles``, which
runs ``_collide_native.capsule_triangles`` (``_collide_native.pyx:454-574``)
one triangle per thread, FP64 and bit-exact with the Cython loop.
"""
import ctypes
from typing import Tuple

import numpy as np

from . import _ffi
from ._ffi import _d, _u8, check, ptr


class CapsuleProxy:
    """A segment plus radius along the local Y axis (reference ``body.py:80-108``)."""

    kind = 'capsule'

    def __init__(self, center, half_height: float, radius: float, rotation) -> None:
        self.center = np.asarray(center, dtype='d')
        self.half_height = float(half_height)
        self.radius = float(radius)
        self.R = np.asarray(rotation, dtype='d')
        axis = self.R @ np.array([0.0, 1.0, 0.0])
        self.p0 = self.center - axis * self.half_height
        self.p1 = self.center + axis * self.half_height


def capsule_triangle_batch(cap: CapsuleProxy, verts: np.ndarray
                           ) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """``(hit, points, normals, depths)`` for triangles ``verts`` ``(N, 3, 3)``
    (triangle, vertex, axis); normals point from the triangle to the capsule."""
    verts = np.ascontiguousarray(verts, dtype='d').reshape(-1, 3, 3)
    n = len(verts)
    hit = np.zeros(n, dtype=np.uint8)
    points = np.zeros((n, 3))
    normals = np.zeros((n, 3))
    depths = np.zeros(n)
    if n:
        p0 = np.ascontiguousarray(cap.p0, dtype='d')
        p1 = np.ascontiguousarray(cap.p1, dtype='d')
        check(_ffi.load().omi_capsule_triangles(n, ptr(p0, _d), ptr(p1, _d), ctypes.c_double(cap.radius),
                                                ptr(verts, _d), ptr(hit, _u8), ptr(points, _d),
                                                ptr(normals, _d), ptr(depths, _d)))
    return hit.astype(bool), points, normals, depths


def capsule_triangles(cap: CapsuleProxy, v0: np.ndarray, v1: np.ndarray, v2: np.ndarray
                      ) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """:func:`capsule_triangle_batch` with the triangles' vertices as three ``(N, 3)`` arrays."""
    v0 = np.asarray(v0, dtype='d').reshape(-1, 3)
    v1 = np.asarray(v1, dtype='d').reshape(-1, 3)
    v2 = np.asarray(v2, dtype='d').reshape(-1, 3)
    return capsule_triangle_batch(cap, np.stack([v0, v1, v2], axis=1))
