"""The oracle: a CPU restatement of the omi_physics contact pipeline (FP64).

TEST INFRASTRUCTURE ONLY.  ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this package as the checker the HIP
path is compared against.  The product package ``omi_physics_amd`` never
imports it, and nothing here is on the product path.

The C library (``oracle.c``) restates the reference's accelerated path; this
module is a thin ctypes binding plus ``OracleWorld``, which steps one world
given the same flat arrays the HIP backend loads
(``omi_physics_amd.world.export_world_arrays``).

Parity pins (each function's independent checks live in ``tests/``):
  refit, broadphase, box/sphere contacts, GJK/EPA, prepare_and_solve and
  multi-step trajectories are checked against golden vectors written by
  importing the reference's pure-Python path (``tests/golden/make_golden.py``),
  and against closed forms, brute force and invariants.
"""
import ctypes
import os
import subprocess
from typing import Any, Dict, Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, '_build', 'liboracle.so')
_lib: Optional[ctypes.CDLL] = None

_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.POINTER(ctypes.c_int64)
_u8 = ctypes.POINTER(ctypes.c_uint8)


class OWDesc(ctypes.Structure):
    """Mirror of ``ow_desc`` in oracle.h."""
    _fields_ = [
        ('n', ctypes.c_int32),
        ('position', _d), ('orientation', _d), ('linear_velocity', _d), ('angular_velocity', _d),
        ('inv_mass', _d), ('inv_inertia', _d),
        ('gravity_factor', _d), ('linear_damping', _d), ('angular_damping', _d),
        ('quadratic_drag', _d), ('sleep_timer', _d),
        ('motion_type', _i32), ('collider_shape', _i32), ('trigger_shape', _i32),
        ('awake', _u8), ('kind', _i32), ('half', _d),
        ('hull_off', _i32), ('hull_cnt', _i32), ('hull_pts', _d),
        ('mat_slot', _i32), ('n_mat', ctypes.c_int32),
        ('friction_table', _d), ('restitution_table', _d),
        ('filt_slot', _i32), ('n_filt', ctypes.c_int32), ('filter_table', _u8),
        ('gravity', ctypes.c_double * 3),
        ('default_linear_damping', ctypes.c_double), ('default_angular_damping', ctypes.c_double),
        ('sleep_enabled', ctypes.c_int32), ('velocity_iterations', ctypes.c_int32),
        ('has_collision', ctypes.c_int32),
        ('n_joints', ctypes.c_int32), ('joint_iterations', ctypes.c_int32),
        ('joint_a', _i32), ('joint_b', _i32), ('joint_anchor', _d),
        ('joint_local_a', _d), ('joint_local_b', _d), ('joint_fixed', _u8),
    ]


def build() -> str:
    """Compile ``liboracle.so`` with the committed Makefile; return its path."""
    subprocess.run(['make', '-s', '-C', _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    """The loaded oracle library (built on first use if absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ow_create.restype = ctypes.c_void_p
        L.ow_create.argtypes = [ctypes.POINTER(OWDesc)]
        L.ow_destroy.argtypes = [ctypes.c_void_p]
        L.ow_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.ow_body_count.argtypes = [ctypes.c_void_p]
        L.ow_get_state.argtypes = [ctypes.c_void_p, _d, _d, _d, _d, _u8, _d, _d, _d]
        L.ow_set_velocity.argtypes = [ctypes.c_void_p, _d, _d]
        L.ow_set_params.argtypes = [ctypes.c_void_p, _d, _d, _d, _d, _d, _d, _d, _d, ctypes.c_int32, ctypes.c_int32]
        L.ow_set_state.argtypes = [ctypes.c_void_p, _d, _d, _d, _d, _u8, _d]
        L.ow_num_contacts.argtypes = [ctypes.c_void_p]
        L.ow_get_contacts.argtypes = [ctypes.c_void_p, _i32, _i32, _d, _d, _d, _d, _d]
        L.ow_num_pairs.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.ow_get_pairs.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i32]
        L.ow_counters.argtypes = [ctypes.c_void_p, _i64]
        L.ow_num_trigger_overlaps.restype = ctypes.c_int64
        L.ow_num_trigger_overlaps.argtypes = [ctypes.c_void_p]
        L.ow_get_trigger_overlaps.argtypes = [ctypes.c_void_p, _i32]
        L.ow_quat_to_matrix.argtypes = [_d, _d]
        L.ow_refit.argtypes = [ctypes.c_int32, _d, _d, _i32, _d, _i32, _i32, _d, ctypes.c_double, _d, _d]
        L.ow_sap_pairs.restype = ctypes.c_int64
        L.ow_sap_pairs.argtypes = [ctypes.c_int32, _d, _d, _i32, _i32, ctypes.c_int32, _u8, _i32,
                                   ctypes.c_int64, _i64]
        L.ow_box_box.restype = ctypes.c_int32
        L.ow_box_box.argtypes = [_d, _d, _d, _d, _d, _d, _d, _d, _d]
        L.ow_sphere_sphere.restype = ctypes.c_int32
        L.ow_sphere_sphere.argtypes = [_d, ctypes.c_double, _d, ctypes.c_double, _d, _d, _d]
        L.ow_sphere_box.restype = ctypes.c_int32
        L.ow_sphere_box.argtypes = [_d, ctypes.c_double, _d, _d, _d, _d, _d, _d]
        L.ow_collide_convex.restype = ctypes.c_int32
        L.ow_collide_convex.argtypes = [ctypes.c_int32, _d, _d, _d, _d, ctypes.c_int32,
                                        ctypes.c_int32, _d, _d, _d, _d, ctypes.c_int32, _d, _d, _d]
        L.ow_gjk_intersect.restype = ctypes.c_int32
        L.ow_gjk_intersect.argtypes = [ctypes.c_int32, _d, _d, _d, _d, ctypes.c_int32,
                                       ctypes.c_int32, _d, _d, _d, _d, ctypes.c_int32]
        L.ow_prepare_and_solve.argtypes = [ctypes.c_int64, _i64, _i64, _d, _d, _d, _d, _d, _d, _d, _d,
                                           _d, _d, _d, _d, ctypes.c_double, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.c_double, ctypes.c_double]
        L.ow_integrate_forces.argtypes = [ctypes.c_int32, _i32, _u8, _d, ctypes.c_int32, _d, _d, _d, _d,
                                          ctypes.c_double, ctypes.c_double, _d, _d, ctypes.c_double]
        L.ow_update_sleep.argtypes = [ctypes.c_int32, ctypes.c_int64, _i32, _i32, _i32, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double, _d, _d, _u8, _d]
        L.ow_integrate_positions.argtypes = [ctypes.c_int32, _i32, _u8, _d, _d, _d, _d, ctypes.c_double]
        L.ow_solve3.argtypes = [_d, _d, _d]
        L.ow_snapshot.argtypes = [ctypes.c_int32, _d, _d, _u8, _i32, _d, _d, _u8, _u8]
        L.ow_capsule_triangles.argtypes = [ctypes.c_int64, _d, _d, ctypes.c_double, _d, _u8, _d, _d, _d]
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t) if a is not None else None


def _c(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


class OracleWorld:
    """One world stepped by the oracle, loaded from ``export_world_arrays`` output."""

    def __init__(self, arrays: Dict[str, Any]) -> None:
        L = lib()
        keep = {}
        def f64(name): keep[name] = _c(arrays[name], np.float64); return _p(keep[name], _d)
        def i32(name): keep[name] = _c(arrays[name], np.int32); return _p(keep[name], _i32)
        def u8(name): keep[name] = _c(arrays[name], np.uint8); return _p(keep[name], _u8)
        d = OWDesc()
        d.n = int(arrays['n'])
        d.position = f64('position'); d.orientation = f64('orientation')
        d.linear_velocity = f64('linear_velocity'); d.angular_velocity = f64('angular_velocity')
        d.inv_mass = f64('inv_mass'); d.inv_inertia = f64('inv_inertia')
        d.gravity_factor = f64('gravity_factor'); d.linear_damping = f64('linear_damping')
        d.angular_damping = f64('angular_damping'); d.quadratic_drag = f64('quadratic_drag')
        d.sleep_timer = f64('sleep_timer')
        d.motion_type = i32('motion_type'); d.collider_shape = i32('collider_shape')
        d.trigger_shape = i32('trigger_shape'); d.awake = u8('awake')
        d.kind = i32('kind'); d.half = f64('half')
        d.hull_off = i32('hull_off'); d.hull_cnt = i32('hull_cnt')
        keep['hull_pts'] = _c(np.asarray(arrays['hull_pts']).reshape(-1, 3), np.float64)
        d.hull_pts = _p(keep['hull_pts'], _d)
        d.mat_slot = i32('mat_slot'); d.n_mat = int(arrays['n_mat'])
        d.friction_table = f64('friction_table'); d.restitution_table = f64('restitution_table')
        d.filt_slot = i32('filt_slot'); d.n_filt = int(arrays['n_filt'])
        d.filter_table = u8('filter_table')
        for r in range(3):
            d.gravity[r] = float(arrays['gravity'][r])
        d.default_linear_damping = float(arrays['default_linear_damping'])
        d.default_angular_damping = float(arrays['default_angular_damping'])
        d.sleep_enabled = int(arrays['sleep_enabled'])
        d.velocity_iterations = int(arrays['velocity_iterations'])
        d.has_collision = int(arrays['has_collision'])
        nj = int(arrays.get('n_joints', 0))
        d.n_joints = nj
        d.joint_iterations = int(arrays.get('joint_iterations', 10))
        keep['ja'] = _c(arrays['joint_a'] if nj else np.zeros(1), np.int32)
        keep['jb'] = _c(arrays['joint_b'] if nj else np.zeros(1), np.int32)
        keep['janc'] = _c(np.asarray(arrays['joint_anchor']).reshape(-1, 3) if nj else np.zeros((1, 3)), np.float64)
        d.joint_a = _p(keep['ja'], _i32); d.joint_b = _p(keep['jb'], _i32); d.joint_anchor = _p(keep['janc'], _d)
        if nj and arrays.get('joint_fixed') is not None:
            keep['jla'] = _c(np.asarray(arrays['joint_local_a']).reshape(-1, 3), np.float64)
            keep['jlb'] = _c(np.asarray(arrays['joint_local_b']).reshape(-1, 3), np.float64)
            keep['jfx'] = _c(arrays['joint_fixed'], np.uint8)
            d.joint_local_a = _p(keep['jla'], _d); d.joint_local_b = _p(keep['jlb'], _d)
            d.joint_fixed = _p(keep['jfx'], _u8)
        self._keep = keep
        self._h = L.ow_create(ctypes.byref(d))
        self.n = d.n

    def __del__(self):
        if getattr(self, '_h', None) and _lib is not None:
            _lib.ow_destroy(self._h)
            self._h = None

    def step(self, dt: float, steps: int = 1) -> None:
        """Advance ``steps`` fixed steps of ``dt``."""
        for _ in range(steps):
            lib().ow_step(self._h, dt)

    def state(self) -> Dict[str, np.ndarray]:
        """Copies of the mutable columns."""
        n = self.n
        out = dict(position=np.empty((n, 3)), orientation=np.empty((n, 4)),
                   linear_velocity=np.empty((n, 3)), angular_velocity=np.empty((n, 3)),
                   awake=np.empty(n, np.uint8), sleep_timer=np.empty(n),
                   aabb_min=np.empty((n, 3)), aabb_max=np.empty((n, 3)))
        lib().ow_get_state(self._h, _p(out['position'], _d), _p(out['orientation'], _d),
                           _p(out['linear_velocity'], _d), _p(out['angular_velocity'], _d),
                           _p(out['awake'], _u8), _p(out['sleep_timer'], _d),
                           _p(out['aabb_min'], _d), _p(out['aabb_max'], _d))
        return out

    def set_velocity(self, lv: np.ndarray, av: np.ndarray) -> None:
        lv = _c(lv, np.float64); av = _c(av, np.float64)
        lib().ow_set_velocity(self._h, _p(lv, _d), _p(av, _d))

    def set_params(self, inv_mass=None, inv_inertia=None, gravity_factor=None, linear_damping=None,
                   angular_damping=None, quadratic_drag=None, gravity=None, default_damping=None,
                   sleep_enabled=None, velocity_iterations=None) -> None:
        """Edit per-body parameters and world settings between steps (None keeps a value)."""
        arrs = [None if x is None else _c(x, np.float64) for x in (inv_mass, inv_inertia, gravity_factor,
                                                                   linear_damping, angular_damping,
                                                                   quadratic_drag, gravity, default_damping)]
        self._keep['params'] = arrs
        lib().ow_set_params(self._h, *[_p(a, _d) for a in arrs],
                            -1 if sleep_enabled is None else int(bool(sleep_enabled)),
                            -1 if velocity_iterations is None else int(velocity_iterations))

    def set_state(self, position=None, orientation=None, linear_velocity=None, angular_velocity=None,
                  awake=None, sleep_timer=None) -> None:
        """Edit the mutable state between steps (None keeps a column)."""
        arrs = [None if x is None else _c(x, np.float64) for x in (position, orientation, linear_velocity,
                                                                   angular_velocity)]
        aw = None if awake is None else _c(awake, np.uint8)
        tm = None if sleep_timer is None else _c(sleep_timer, np.float64)
        lib().ow_set_state(self._h, *[_p(a, _d) for a in arrs], _p(aw, _u8), _p(tm, _d))

    def contacts(self) -> Dict[str, np.ndarray]:
        """The last step's contacts in solver order, with their final impulses."""
        K = lib().ow_num_contacts(self._h)
        out = dict(a=np.empty(K, np.int32), b=np.empty(K, np.int32), point=np.empty((K, 3)),
                   normal=np.empty((K, 3)), depth=np.empty(K), normal_impulse=np.empty(K),
                   tangent_impulse=np.empty((K, 2)))
        lib().ow_get_contacts(self._h, _p(out['a'], _i32), _p(out['b'], _i32),
                              _p(out['point'], _d), _p(out['normal'], _d), _p(out['depth'], _d),
                              _p(out['normal_impulse'], _d), _p(out['tangent_impulse'], _d))
        return out

    def pairs(self, solid: bool = False) -> np.ndarray:
        """The last step's broadphase pairs (or the solid subset), ``(P, 2)``."""
        P = lib().ow_num_pairs(self._h, int(solid))
        out = np.empty((P, 2), np.int32)
        if P:
            lib().ow_get_pairs(self._h, int(solid), _p(out, _i32))
        return out

    def trigger_overlaps(self) -> np.ndarray:
        """The last step's trigger overlaps ``(k, 2)``: (trigger body, other body), sorted."""
        k = int(lib().ow_num_trigger_overlaps(self._h))
        out = np.empty((k, 2), np.int32)
        if k:
            lib().ow_get_trigger_overlaps(self._h, _p(out, _i32))
            out = out[np.lexsort((out[:, 1], out[:, 0]))]
        return out

    def counters(self) -> Dict[str, int]:
        o = np.zeros(5, np.int64)
        lib().ow_counters(self._h, _p(o, _i64))
        return dict(pairs=int(o[0]), solid_pairs=int(o[1]), x_candidates=int(o[2]),
                    manifolds=int(o[3]), contacts=int(o[4]))


# ---- stage-level functions --------------------------------------------------

def quat_to_matrix(q) -> np.ndarray:
    q = _c(q, np.float64); R = np.empty(9)
    lib().ow_quat_to_matrix(_p(q, _d), _p(R, _d))
    return R.reshape(3, 3)


def capsule_triangles(p0, p1, radius, verts):
    """``(hit, points, normals, depths)`` of the capsule (segment p0 p1, radius) against
    triangles ``verts`` ``(n, 3, 3)``."""
    p0 = _c(p0, np.float64); p1 = _c(p1, np.float64)
    verts = _c(np.asarray(verts, dtype=np.float64).reshape(-1, 3, 3), np.float64); n = len(verts)
    hit = np.zeros(n, np.uint8); pts = np.zeros((n, 3)); nrm = np.zeros((n, 3)); dep = np.zeros(n)
    lib().ow_capsule_triangles(n, _p(p0, _d), _p(p1, _d), float(radius), _p(verts, _d), _p(hit, _u8),
                               _p(pts, _d), _p(nrm, _d), _p(dep, _d))
    return hit.astype(bool), pts, nrm, dep


def snapshot(pos, quat, awake, motion_type):
    """``(position, axis_angle, awake, dynamic)`` of n bodies (the reference's snapshot)."""
    pos = _c(pos, np.float64); quat = _c(quat, np.float64)
    aw = _c(awake, np.uint8); mt = _c(motion_type, np.int32); n = len(pos)
    op = np.empty((n, 3)); aa = np.empty((n, 4)); oa = np.empty(n, np.uint8); dy = np.empty(n, np.uint8)
    lib().ow_snapshot(n, _p(pos, _d), _p(quat, _d), _p(aw, _u8), _p(mt, _i32), _p(op, _d), _p(aa, _d),
                      _p(oa, _u8), _p(dy, _u8))
    return op, aa, oa.astype(bool), dy.astype(bool)


def refit(pos, quat, kind, half, margin=0.05, hull_off=None, hull_cnt=None, hull_pts=None):
    pos = _c(pos, np.float64); quat = _c(quat, np.float64); kind = _c(kind, np.int32)
    half = _c(half, np.float64); n = len(pos)
    hull_off = _c(hull_off if hull_off is not None else np.full(n, -1), np.int32)
    hull_cnt = _c(hull_cnt if hull_cnt is not None else np.zeros(n), np.int32)
    hull_pts = _c(hull_pts if hull_pts is not None else np.zeros((1, 3)), np.float64)
    lo = np.empty((n, 3)); hi = np.empty((n, 3))
    lib().ow_refit(n, _p(pos, _d), _p(quat, _d), _p(kind, _i32), _p(half, _d), _p(hull_off, _i32),
                   _p(hull_cnt, _i32), _p(hull_pts, _d), float(margin), _p(lo, _d), _p(hi, _d))
    return lo, hi


def sap_pairs(lo, hi, motion_type, filt_slot=None, filter_table=None) -> Tuple[np.ndarray, int]:
    lo = _c(lo, np.float64); hi = _c(hi, np.float64); mt = _c(motion_type, np.int32); n = len(lo)
    fs = _c(filt_slot if filt_slot is not None else np.zeros(n), np.int32)
    ft = _c(filter_table if filter_table is not None else np.ones((1, 1)), np.uint8)
    nf = int(round(np.sqrt(ft.size)))
    xc = np.zeros(1, np.int64)
    cap = 1024
    while True:
        out = np.empty((cap, 2), np.int32)
        np_ = lib().ow_sap_pairs(n, _p(lo, _d), _p(hi, _d), _p(mt, _i32), _p(fs, _i32), nf,
                                 _p(ft, _u8), _p(out, _i32), cap, _p(xc, _i64))
        if np_ <= cap:
            return out[:np_].copy(), int(xc[0])
        cap = int(np_)


def box_box(ca, Ra, ha, cb, Rb, hb):
    args = [_c(x, np.float64) for x in (ca, Ra, ha, cb, Rb, hb)]
    nrm = np.zeros(3); pts = np.zeros((4, 3)); dep = np.zeros(4)
    n = lib().ow_box_box(*[_p(a, _d) for a in args], _p(nrm, _d), _p(pts, _d), _p(dep, _d))
    return n, nrm, pts[:n], dep[:n]


def sphere_sphere(ca, ra, cb, rb):
    ca = _c(ca, np.float64); cb = _c(cb, np.float64)
    nrm = np.zeros(3); pt = np.zeros(3); dep = np.zeros(1)
    n = lib().ow_sphere_sphere(_p(ca, _d), float(ra), _p(cb, _d), float(rb), _p(nrm, _d), _p(pt, _d), _p(dep, _d))
    return n, nrm, pt[None, :][:n], dep[:n]


def sphere_box(cs, rs, cb, Rb, hb):
    cs = _c(cs, np.float64); cb = _c(cb, np.float64); Rb = _c(Rb, np.float64); hb = _c(hb, np.float64)
    nrm = np.zeros(3); pt = np.zeros(3); dep = np.zeros(1)
    n = lib().ow_sphere_box(_p(cs, _d), float(rs), _p(cb, _d), _p(Rb, _d), _p(hb, _d), _p(nrm, _d), _p(pt, _d), _p(dep, _d))
    return n, nrm, pt[None, :][:n], dep[:n]


def _proxy_args(kind, c, R, h, pts):
    c = _c(c, np.float64); R = _c(R, np.float64)
    h = _c(h if h is not None else np.zeros(3), np.float64)
    pts = _c(pts if pts is not None else np.zeros((1, 3)), np.float64)
    m = len(pts) if kind == 3 else 0
    return (int(kind), _p(c, _d), _p(R, _d), _p(h, _d), _p(pts, _d), m), (c, R, h, pts)


def collide_convex(ka, ca, Ra, ha, pts_a, kb, cb, Rb, hb, pts_b):
    A, ka_ = _proxy_args(ka, ca, Ra, ha, pts_a)
    B, kb_ = _proxy_args(kb, cb, Rb, hb, pts_b)
    nrm = np.zeros(3); pt = np.zeros(3); dep = np.zeros(1)
    n = lib().ow_collide_convex(*A, *B, _p(nrm, _d), _p(pt, _d), _p(dep, _d))
    return n, nrm, pt, float(dep[0])


def gjk_intersect(ka, ca, Ra, ha, pts_a, kb, cb, Rb, hb, pts_b) -> bool:
    A, ka_ = _proxy_args(ka, ca, Ra, ha, pts_a)
    B, kb_ = _proxy_args(kb, cb, Rb, hb, pts_b)
    return bool(lib().ow_gjk_intersect(*A, *B))


def solve3(K, rhs) -> np.ndarray:
    K = _c(K, np.float64); rhs = _c(rhs, np.float64); x = np.zeros(3)
    lib().ow_solve3(_p(K, _d), _p(rhs, _d), _p(x, _d))
    return x


def integrate_forces(motion_type, awake, gravity, gravity_factor, linear_damping, angular_damping,
                     quadratic_drag, default_linear_damping, default_angular_damping, lv, av, dt) -> None:
    """In place on ``lv, av`` (``(n, 3)``); ``gravity`` is ``(3,)`` or per body ``(n, 3)``."""
    mt = _c(motion_type, np.int32); aw = _c(awake, np.uint8); n = len(mt)
    g = _c(gravity, np.float64)
    cols = [_c(x, np.float64) for x in (gravity_factor, linear_damping, angular_damping, quadratic_drag)]
    for x in (lv, av):
        assert x.flags['C_CONTIGUOUS'] and x.dtype == np.float64
    lib().ow_integrate_forces(n, _p(mt, _i32), _p(aw, _u8), _p(g, _d), 3 if g.ndim == 2 else 0,
                              *[_p(x, _d) for x in cols], float(default_linear_damping),
                              float(default_angular_damping), _p(lv, _d), _p(av, _d), float(dt))


def integrate_positions(motion_type, awake, lv, av, pos, quat, dt) -> None:
    """In place on ``pos`` ``(n, 3)`` and ``quat`` ``(n, 4)``."""
    mt = _c(motion_type, np.int32); aw = _c(awake, np.uint8); n = len(mt)
    lv = _c(lv, np.float64); av = _c(av, np.float64)
    for x in (pos, quat):
        assert x.flags['C_CONTIGUOUS'] and x.dtype == np.float64
    lib().ow_integrate_positions(n, _p(mt, _i32), _p(aw, _u8), _p(lv, _d), _p(av, _d), _p(pos, _d),
                                 _p(quat, _d), float(dt))


def update_sleep(ca, cb, motion_type, lv, av, awake, sleep_timer, dt, lin=0.08, ang=0.25, t_sleep=0.4) -> None:
    """In place on ``lv, av, awake (uint8), sleep_timer`` (the step's contacts ``ca, cb``)."""
    ca = _c(ca, np.int32); cb = _c(cb, np.int32); mt = _c(motion_type, np.int32)
    for x, t in ((lv, np.float64), (av, np.float64), (awake, np.uint8), (sleep_timer, np.float64)):
        assert x.flags['C_CONTIGUOUS'] and x.dtype == t
    lib().ow_update_sleep(len(mt), len(ca), _p(ca, _i32), _p(cb, _i32), _p(mt, _i32), float(dt), float(lin),
                          float(ang), float(t_sleep), _p(lv, _d), _p(av, _d), _p(awake, _u8), _p(sleep_timer, _d))


def prepare_and_solve(a_idx, b_idx, point, normal, depth, pos, lv, av, inv_mass, invIw,
                      restitution, friction, nImp, tImp, restitution_threshold=0.5,
                      iters=10, warm_start=True, correction=0.8, slop=0.005) -> None:
    """In-place on ``pos, lv, av, nImp, tImp`` like the reference entry point."""
    a = _c(a_idx, np.int64); b = _c(b_idx, np.int64)
    arrs = [_c(x, np.float64) for x in (point, normal, depth, inv_mass, invIw, restitution, friction)]
    for x in (pos, lv, av, nImp, tImp):
        assert x.flags['C_CONTIGUOUS'] and x.dtype == np.float64
    lib().ow_prepare_and_solve(len(a), _p(a, _i64), _p(b, _i64), _p(arrs[0], _d), _p(arrs[1], _d),
                               _p(arrs[2], _d), _p(pos, _d), _p(lv, _d), _p(av, _d), _p(arrs[3], _d),
                               _p(arrs[4], _d), _p(arrs[5], _d), _p(arrs[6], _d), _p(nImp, _d),
                               _p(tImp, _d), float(restitution_threshold), int(iters),
                               int(bool(warm_start)), float(correction), float(slop))
