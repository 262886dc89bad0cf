"""Host-side quaternion helpers, ``(x, y, z, w)`` order (reference ``mathutil.py``).

These run on the host only (scene set-up, render writeback, tests).  The device
kernels carry their own FP64 restatement of the same formulas with the same
operation order, so a pose computed here and one computed on the GPU agree to
the last bit.
"""
from typing import Sequence, Union

import numpy as np

Vec = Union[Sequence[float], np.ndarray]


def normalize(v: Vec, axis: int = -1, eps: float = 1e-12) -> np.ndarray:
    """``v / max(|v|, eps)`` along ``axis`` (reference ``mathutil.py:21-24``)."""
    v = np.asarray(v, dtype='d')
    n = np.linalg.norm(v, axis=axis, keepdims=True)
    return v / np.maximum(n, eps)


def quat_normalize(q: Vec) -> np.ndarray:
    """Unit quaternion(s)."""
    return normalize(q, axis=-1)


def quat_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Hamilton product ``a ⊗ b`` (reference ``mathutil.py:32-41``)."""
    ax, ay, az, aw = a[..., 0], a[..., 1], a[..., 2], a[..., 3]
    bx, by, bz, bw = b[..., 0], b[..., 1], b[..., 2], b[..., 3]
    return np.stack([
        aw * bx + ax * bw + ay * bz - az * by,
        aw * by - ax * bz + ay * bw + az * bx,
        aw * bz + ax * by - ay * bx + az * bw,
        aw * bw - ax * bx - ay * by - az * bz,
    ], axis=-1)


def quat_to_matrix(q: np.ndarray) -> np.ndarray:
    """``(..., 3, 3)`` rotation matrices, columns = body axes (reference ``mathutil.py:59-76``)."""
    q = quat_normalize(q)
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    xx, yy, zz = x * x, y * y, z * z
    xy, xz, yz = x * y, x * z, y * z
    wx, wy, wz = w * x, w * y, w * z
    m = np.empty(q.shape[:-1] + (3, 3), dtype=q.dtype)
    m[..., 0, 0] = 1 - 2 * (yy + zz)
    m[..., 0, 1] = 2 * (xy - wz)
    m[..., 0, 2] = 2 * (xz + wy)
    m[..., 1, 0] = 2 * (xy + wz)
    m[..., 1, 1] = 1 - 2 * (xx + zz)
    m[..., 1, 2] = 2 * (yz - wx)
    m[..., 2, 0] = 2 * (xz - wy)
    m[..., 2, 1] = 2 * (yz + wx)
    m[..., 2, 2] = 1 - 2 * (xx + yy)
    return m


def quat_integrate(q: np.ndarray, omega: np.ndarray, dt: float) -> np.ndarray:
    """``normalize(q + ½·dt·(ω,0)⊗q)`` (reference ``mathutil.py:108-113``)."""
    w = np.zeros(q.shape, dtype=q.dtype)
    w[..., :3] = omega
    dq = 0.5 * quat_mul(w, q) * dt
    return quat_normalize(q + dq)


def quat_from_axis_angle(axis: Vec, angle: float) -> np.ndarray:
    """Unit quaternion rotating ``angle`` radians about ``axis``."""
    axis = normalize(np.asarray(axis, dtype='d'))
    half = 0.5 * angle
    s = np.sin(half)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, np.cos(half)])


def quat_to_axis_angle(q: np.ndarray) -> np.ndarray:
    """Quaternion(s) -> ``(x, y, z, angle)``; near-identity gives ``(0, 1, 0, 0)``."""
    q = quat_normalize(q)
    w = np.clip(q[..., 3], -1.0, 1.0)
    s = np.sqrt(np.maximum(1.0 - w * w, 0.0))
    small = s < 1e-9
    axis = q[..., :3] / np.where(small, 1.0, s)[..., None]
    out = np.empty(q.shape[:-1] + (4,), dtype=q.dtype)
    out[..., :3] = np.where(small[..., None], np.array([0.0, 1.0, 0.0]), axis)
    out[..., 3] = np.where(small, 0.0, 2.0 * np.arccos(w))
    return out
