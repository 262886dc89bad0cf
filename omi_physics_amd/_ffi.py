"""ctypes binding of ``libomi_hip.so`` (declared in ``include/omi_hip.h``).

The library is built in-tree by ``omi_physics_amd/csrc/Makefile`` into
``omi_physics_amd/_lib/``.  :func:`load` raises ``RuntimeError`` when it is
missing: there is no CPU substitute for the HIP contact pipeline.
"""
import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, '_lib', 'libomi_hip.so')
CSRC = os.path.join(HERE, 'csrc')

OMI_FP64, OMI_FP32 = 64, 32
OMI_OK, OMI_ERR_ARG, OMI_ERR_HIP, OMI_ERR_CAPACITY, OMI_ERR_STATE = 0, -1, -2, -3, -4

_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.POINTER(ctypes.c_int64)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p


class WorldDesc(ctypes.Structure):
    """Mirror of ``omi_world_desc``."""
    _fields_ = [
        ('n', ctypes.c_int32),
        ('position', _d), ('orientation', _d), ('linear_velocity', _d), ('angular_velocity', _d),
        ('prev_position', _d), ('prev_orientation', _d),
        ('inv_mass', _d), ('inv_inertia', _d),
        ('gravity_factor', _d), ('linear_damping', _d), ('angular_damping', _d),
        ('quadratic_drag', _d), ('sleep_timer', _d),
        ('motion_type', _i32), ('collider_shape', _i32), ('trigger_shape', _i32),
        ('awake', _u8), ('kind', _i32), ('half', _d),
        ('hull_off', _i32), ('hull_cnt', _i32), ('hull_pts', _d), ('n_hull_pts', ctypes.c_int64),
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


class BatchConfig(ctypes.Structure):
    """Mirror of ``omi_batch_config``."""
    _fields_ = [
        ('max_worlds', ctypes.c_int32), ('precision', ctypes.c_int32),
        ('max_bodies', ctypes.c_int64), ('max_pairs', ctypes.c_int64),
        ('max_contacts', ctypes.c_int64), ('max_hull_pts', ctypes.c_int64), ('max_joints', ctypes.c_int64),
        ('record_broadphase', ctypes.c_int32), ('solver_block', ctypes.c_int32),
        ('sap_block', ctypes.c_int32 * 4), ('solve_block', ctypes.c_int32 * 4),
        ('max_candidates', ctypes.c_int64), ('max_solver_iterations', ctypes.c_int32),
        ('max_trigger_pairs', ctypes.c_int64), ('island_solve', ctypes.c_int32), ('clip16_all', ctypes.c_int32),
    ]


#: every symbol include/omi_hip.h declares
EXPORTS = (
    'omi_last_error', 'omi_abi_version', 'omi_device_count', 'omi_device_memory',
    'omi_refit', 'omi_sap_pairs', 'omi_box_box', 'omi_sphere_sphere', 'omi_sphere_box',
    'omi_collide_convex', 'omi_collide_convex_noalloc', 'omi_capsule_triangles', 'omi_prepare_and_solve',
    'omi_batch_create', 'omi_batch_destroy', 'omi_batch_load_world', 'omi_batch_remove_world',
    'omi_integrate_forces', 'omi_integrate_positions', 'omi_sleep',
    'omi_batch_set_state', 'omi_batch_set_aabbs', 'omi_batch_set_params', 'omi_batch_read_cache',
    'omi_batch_write_cache', 'omi_batch_collide', 'omi_batch_read_joint_anchors',
    'omi_batch_prepare', 'omi_batch_step', 'omi_batch_refit', 'omi_batch_sync',
    'omi_batch_read_world', 'omi_batch_fetch_world', 'omi_batch_fetch_ready', 'omi_batch_fetch_wait',
    'omi_batch_mark', 'omi_batch_mark_wait',
    'omi_batch_snapshot', 'omi_batch_read_contacts', 'omi_batch_read_pairs',
    'omi_batch_read_trigger_overlaps',
    'omi_batch_counters', 'omi_batch_set_timing', 'omi_batch_kernel_times',
    'omi_batch_kernel_names', 'omi_batch_use_graph', 'omi_batch_set_blocks', 'omi_batch_world_stats',
    'omi_batch_islands',
    'omi_batch_device_bytes',
    'omi_group_create', 'omi_group_destroy', 'omi_group_prepare', 'omi_group_step', 'omi_group_sync',
)

_lib: Optional[ctypes.CDLL] = None


def build(arch: str = 'gfx950') -> str:
    """Compile the HIP library for ``arch`` with the in-tree Makefile (works without a GPU)."""
    subprocess.run(['make', '-s', '-C', CSRC, 'ARCH=%s' % arch], check=True)
    return LIB_PATH


def load() -> ctypes.CDLL:
    """Load ``libomi_hip.so``; raise ``RuntimeError`` if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError('libomi_hip.so is not built (%s); run __graft_entry__.build() or '
                           'make -C omi_physics_amd/csrc' % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        'omi_last_error': (ctypes.c_char_p, []),
        'omi_abi_version': (ctypes.c_int32, []),
        'omi_device_count': (ctypes.c_int32, []),
        'omi_device_memory': (ctypes.c_int32, [_i64, _i64]),
        'omi_refit': (ctypes.c_int32, [ctypes.c_int32, _d, _d, _i32, _d, _i32, _i32, _d, ctypes.c_int64,
                                       ctypes.c_double, _d, _d]),
        'omi_sap_pairs': (ctypes.c_int32, [ctypes.c_int32, _d, _d, _i32, _i32, ctypes.c_int32, _u8, _i32,
                                           ctypes.c_int64, _i64]),
        'omi_box_box': (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32, _d, _d, _d, _d, _d, _d, _i32, _d, _d, _d]),
        'omi_sphere_sphere': (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32, _d, _d, _d, _d, _i32, _d, _d, _d]),
        'omi_sphere_box': (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32, _d, _d, _d, _d, _d, _i32, _d, _d, _d]),
        'omi_collide_convex': (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32, _i32, _d, _d, _d, _i32, _i32,
                                                _i32, _d, _d, _d, _i32, _i32, _d, ctypes.c_int64,
                                                _i32, _d, _d, _d]),
        'omi_collide_convex_noalloc': (ctypes.c_int32, [ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                                        _vp, _vp, _vp, _vp, ctypes.c_int64, _vp, _vp, _vp, _vp,
                                                        _vp]),
        'omi_capsule_triangles': (ctypes.c_int32, [ctypes.c_int64, _d, _d, ctypes.c_double, _d, _u8, _d, _d, _d]),
        'omi_prepare_and_solve': (ctypes.c_int32, [ctypes.c_int64, ctypes.c_int32, _i64, _i64, _d, _d, _d, _d, _d,
                                                   _d, _d, _d, _d, _d, _d, _d, ctypes.c_double, ctypes.c_int32,
                                                   ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_int32]),
        'omi_batch_create': (ctypes.c_int32, [ctypes.POINTER(BatchConfig), ctypes.POINTER(_vp)]),
        'omi_batch_destroy': (ctypes.c_int32, [_vp]),
        'omi_batch_load_world': (ctypes.c_int32, [_vp, ctypes.c_int32, ctypes.POINTER(WorldDesc)]),
        'omi_batch_remove_world': (ctypes.c_int32, [_vp, ctypes.c_int32]),
        'omi_batch_set_state': (ctypes.c_int32, [_vp, ctypes.c_int32, _d, _d, _d, _d, _u8, _d]),
        'omi_batch_set_aabbs': (ctypes.c_int32, [_vp, ctypes.c_int32, _d, _d]),
        'omi_batch_read_joint_anchors': (ctypes.c_int32, [_vp, ctypes.c_int32, _d, _d, _u8]),
        'omi_batch_set_params': (ctypes.c_int32, [_vp, ctypes.c_int32, _d, _d, _d, _d, _d, _d, _d, _d,
                                                  ctypes.c_int32, ctypes.c_int32]),
        'omi_batch_read_cache': (ctypes.c_int32, [_vp, ctypes.c_int32, ctypes.c_int64, _i64, _i32, _i32, _d, _d,
                                                  _d]),
        'omi_batch_write_cache': (ctypes.c_int32, [_vp, ctypes.c_int32, ctypes.c_int64, _i32, _i32, _d, _d, _d]),
        'omi_batch_collide': (ctypes.c_int32, [_vp, ctypes.c_double]),
        'omi_integrate_forces': (ctypes.c_int32, [ctypes.c_int32, _i32, _u8, _d, _d, _d, _d, _d, ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_double, _d, _d]),
        'omi_integrate_positions': (ctypes.c_int32, [ctypes.c_int32, _i32, _u8, _d, _d, ctypes.c_double, _d, _d]),
        'omi_sleep': (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int64, _i32, _i32, _i32, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, _d, _d, _u8, _d]),
        'omi_batch_prepare': (ctypes.c_int32, [_vp, ctypes.c_double]),
        'omi_batch_step': (ctypes.c_int32, [_vp, ctypes.c_double, ctypes.c_int32]),
        'omi_batch_refit': (ctypes.c_int32, [_vp, ctypes.c_double]),
        'omi_batch_sync': (ctypes.c_int32, [_vp]),
        'omi_batch_read_world': (ctypes.c_int32, [_vp, ctypes.c_int32, _d, _d, _d, _d, _u8, _d, _d, _d, _d, _d]),
        'omi_batch_fetch_world': (ctypes.c_int32, [_vp, ctypes.c_int32, _i64]),
        'omi_batch_fetch_ready': (ctypes.c_int32, [_vp, ctypes.c_int64]),
        'omi_batch_mark': (ctypes.c_int32, [_vp, _i64]),
        'omi_batch_mark_wait': (ctypes.c_int32, [_vp, ctypes.c_int64]),
        'omi_batch_fetch_wait': (ctypes.c_int32, [_vp, ctypes.c_int64, _d, _d, _d, _d, _u8, _d, _d, _d, _d, _d]),
        'omi_batch_snapshot': (ctypes.c_int32, [_vp, ctypes.c_int32, _i32, _i64, _d, _d, _u8, _u8]),
        'omi_batch_read_trigger_overlaps': (ctypes.c_int32, [_vp, ctypes.c_int32, ctypes.c_int64, _i64, _i32, _i32]),
        'omi_batch_read_contacts': (ctypes.c_int32, [_vp, ctypes.c_int32, ctypes.c_int64, _i64, _i32, _i32,
                                                     _d, _d, _d, _d, _d]),
        'omi_batch_read_pairs': (ctypes.c_int32, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, _i64,
                                                  _i32]),
        'omi_batch_counters': (ctypes.c_int32, [_vp, _i64]),
        'omi_batch_islands': (ctypes.c_int32, [_vp, _i32, _i64]),
        'omi_batch_set_timing': (ctypes.c_int32, [_vp, ctypes.c_int32]),
        'omi_batch_kernel_times': (ctypes.c_int32, [_vp, ctypes.c_int32, _d]),
        'omi_batch_kernel_names': (ctypes.c_char_p, []),
        'omi_batch_use_graph': (ctypes.c_int32, [_vp, ctypes.c_int32]),
        'omi_batch_set_blocks': (ctypes.c_int32, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
        'omi_batch_world_stats': (ctypes.c_int32, [_vp, ctypes.c_int32] + [ctypes.POINTER(ctypes.c_int32)] * 5),
        'omi_batch_device_bytes': (ctypes.c_int64, [_vp]),
        'omi_group_create': (ctypes.c_int32, [ctypes.POINTER(_vp), ctypes.c_int32, ctypes.POINTER(_vp)]),
        'omi_group_destroy': (ctypes.c_int32, [_vp]),
        'omi_group_prepare': (ctypes.c_int32, [_vp, ctypes.c_double]),
        'omi_group_step': (ctypes.c_int32, [_vp, ctypes.c_double, ctypes.c_int32]),
        'omi_group_sync': (ctypes.c_int32, [_vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    """Raise ``RuntimeError`` with the library's message for a non-zero status."""
    if rc != 0:
        msg = load().omi_last_error().decode(errors='replace')
        raise RuntimeError('omi_hip error %d: %s' % (rc, msg))


def ptr(a: Optional[np.ndarray], t):
    """ctypes pointer to a C-contiguous array (or NULL)."""
    if a is None:
        return None
    assert a.flags['C_CONTIGUOUS'], 'array must be C-contiguous'
    return a.ctypes.data_as(t)


def c64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def ci32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def world_desc(arrays: dict):
    """A ``WorldDesc`` over ``export_world_arrays`` output; returns (desc, keepalive)."""
    keep = {}
    def f64(name):
        keep[name] = c64(arrays[name]); return ptr(keep[name], _d)
    def i32(name):
        keep[name] = ci32(arrays[name]); return ptr(keep[name], _i32)
    d = WorldDesc()
    d.n = int(arrays['n'])
    d.position = f64('position'); d.orientation = f64('orientation')
    d.linear_velocity = f64('linear_velocity'); d.angular_velocity = f64('angular_velocity')
    d.prev_position = f64('prev_position'); d.prev_orientation = f64('prev_orientation')
    d.inv_mass = f64('inv_mass'); d.inv_inertia = f64('inv_inertia')
    d.gravity_factor = f64('gravity_factor'); d.linear_damping = f64('linear_damping')
    d.angular_damping = f64('angular_damping'); d.quadratic_drag = f64('quadratic_drag')
    d.sleep_timer = f64('sleep_timer')
    d.motion_type = i32('motion_type'); d.collider_shape = i32('collider_shape')
    d.trigger_shape = i32('trigger_shape')
    keep['awake'] = np.ascontiguousarray(arrays['awake'], dtype=np.uint8)
    d.awake = ptr(keep['awake'], _u8)
    d.kind = i32('kind'); d.half = f64('half')
    d.hull_off = i32('hull_off'); d.hull_cnt = i32('hull_cnt')
    keep['hull_pts'] = c64(np.asarray(arrays['hull_pts']).reshape(-1, 3))
    if len(keep['hull_pts']) == 0:
        keep['hull_pts'] = np.zeros((1, 3))
        d.n_hull_pts = 0
    else:
        d.n_hull_pts = len(keep['hull_pts'])
    d.hull_pts = ptr(keep['hull_pts'], _d)
    d.mat_slot = i32('mat_slot'); d.n_mat = int(arrays['n_mat'])
    d.friction_table = f64('friction_table'); d.restitution_table = f64('restitution_table')
    d.filt_slot = i32('filt_slot'); d.n_filt = int(arrays['n_filt'])
    keep['filter_table'] = np.ascontiguousarray(arrays['filter_table'], dtype=np.uint8)
    d.filter_table = ptr(keep['filter_table'], _u8)
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
    keep['ja'] = ci32(arrays['joint_a'] if nj else np.zeros(1))
    keep['jb'] = ci32(arrays['joint_b'] if nj else np.zeros(1))
    keep['janc'] = c64(np.asarray(arrays['joint_anchor']).reshape(-1, 3) if nj else np.zeros((1, 3)))
    d.joint_a = ptr(keep['ja'], _i32); d.joint_b = ptr(keep['jb'], _i32)
    d.joint_anchor = ptr(keep['janc'], _d)
    if nj and arrays.get('joint_fixed') is not None:
        keep['jla'] = c64(np.asarray(arrays['joint_local_a']).reshape(-1, 3))
        keep['jlb'] = c64(np.asarray(arrays['joint_local_b']).reshape(-1, 3))
        keep['jfx'] = np.ascontiguousarray(arrays['joint_fixed'], dtype=np.uint8)
        d.joint_local_a = ptr(keep['jla'], _d); d.joint_local_b = ptr(keep['jlb'], _d)
        d.joint_fixed = ptr(keep['jfx'], _u8)
    return d, keep
