"""A world's columns and tables in the layout ``omi_batch_load_world`` loads.

Works on any world object with the reference ``PhysicsWorld``'s public surface
(``world.py:145-223``): the column views (``position`` ... ``sleep_timer``),
``body_count``, ``shapes``, ``materials``, ``filters``, ``material_for``,
``filter_for``, ``pair_friction_for``, ``gravity``, the damping defaults,
``sleep_enabled`` and ``joint_constraints`` — so the reference's own world and
``omi_physics_amd.PhysicsWorld`` load the same way.

Shapes become per-body kind codes, half extents (box) or radius (sphere), and
hull ranges into one point pool that holds each convex shape once.  Material
and filter indices are remapped to dense slots over the values the bodies use,
and the per-slot-pair friction, restitution (``solver.py:252-259``) and filter
outcomes (``broadphase.py:297-301``) are tabulated, so the device reads one
table entry per pair.
"""
from typing import Any, Dict, List, Optional, Tuple

import numpy as np

from . import model

#: shape-kind codes shared with the device and the C ABI
KIND_NONE, KIND_SPHERE, KIND_BOX, KIND_CONVEX = 0, 1, 2, 3
DYNAMIC_CODE = 2


def uniform_gravity(world: Any) -> np.ndarray:
    """The world's base gravity vector: direction normalised, times magnitude
    (``world.py:386-396`` without gravity volumes)."""
    g = world.gravity
    base = np.asarray(g.direction, dtype='d')
    norm = np.linalg.norm(base)
    base = base / norm if norm else base
    return base * g.gravity


def shape_tables(shapes: List[model.Shape]) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """Per shape: kind, half extents / radius, hull offset and count into a point pool; the pool."""
    S = len(shapes)
    kind = np.zeros(S, np.int32)
    half = np.zeros((S, 3))
    off = np.full(S, -1, np.int32)
    cnt = np.zeros(S, np.int32)
    pool: List[np.ndarray] = []
    used = 0
    for s, sh in enumerate(shapes):
        if sh.type == 'sphere':
            kind[s], half[s] = KIND_SPHERE, sh.radius
        elif sh.type == 'box':
            kind[s], half[s] = KIND_BOX, 0.5 * np.asarray(sh.size, dtype='d')
        elif sh.type == 'convex':
            pts = np.ascontiguousarray(sh.points, dtype='d').reshape(-1, 3)
            kind[s], off[s], cnt[s] = KIND_CONVEX, used, len(pts)
            pool.append(pts)
            used += len(pts)
        else:
            kind[s] = -1                      # not built on the device; refused if a body uses it
    return kind, half, off, cnt, (np.concatenate(pool) if pool else np.zeros((0, 3)))


def body_shapes(shapes: List[model.Shape], collider: np.ndarray,
                trigger: np.ndarray) -> Dict[str, np.ndarray]:
    """Per body: the kind, half extents and hull range of its collider shape, else its
    trigger shape (the reference's refit choice, ``world.py:443-466``)."""
    n = len(collider)
    kind_s, half_s, off_s, cnt_s, pool = shape_tables(shapes)
    idx = np.where(collider >= 0, collider, trigger).astype(np.int64)
    both = (collider >= 0) & (trigger >= 0) & (collider != trigger)
    if both.any():
        raise NotImplementedError('body %d has a collider and a different trigger shape (the device keeps one '
                                  'shape per body)' % int(np.flatnonzero(both)[0]))
    has = idx >= 0
    bad = has & (kind_s[np.where(has, idx, 0)] < 0) if len(shapes) else np.zeros(n, bool)
    if bad.any():
        i = int(np.flatnonzero(bad)[0])
        raise ValueError('body %d uses a %r shape: the contact pipeline builds box, sphere and convex shapes'
                         % (i, shapes[int(idx[i])].type))
    kind = np.zeros(n, np.int32)
    half = np.zeros((n, 3))
    off = np.full(n, -1, np.int32)
    cnt = np.zeros(n, np.int32)
    kind[has], half[has], off[has], cnt[has] = kind_s[idx[has]], half_s[idx[has]], off_s[idx[has]], cnt_s[idx[has]]
    return dict(kind=kind, half=half, hull_off=off, hull_cnt=cnt, hull_pts=pool)


def _dense_slots(raw: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    values, slots = np.unique(raw, return_inverse=True)
    return values.astype(np.int64), slots.reshape(-1).astype(np.int32)


def material_tables(world: Any, raw: np.ndarray) -> Dict[str, Any]:
    """Dense material slots and the friction / restitution of every slot pair.

    Restitution combines with body a's rule; friction is the pair override's
    dynamic coefficient, else the dynamic coefficients combined by a's rule
    (``solver.py:252-259``, ``model.py:231-243``)."""
    vals, slot = _dense_slots(raw)
    mats = [world.material_for(int(v)) for v in vals]
    m = len(vals)
    fric, rest = np.zeros((m, m)), np.zeros((m, m))
    for i, ma in enumerate(mats):
        for j, mb in enumerate(mats):
            rest[i, j] = model.combine(ma.restitution, mb.restitution, ma.restitutionCombine)
            override = world.pair_friction_for(int(vals[i]), int(vals[j]))
            fric[i, j] = (override[1] if override is not None
                          else model.combine(ma.dynamicFriction, mb.dynamicFriction, ma.frictionCombine))
    return dict(mat_slot=slot, n_mat=m, friction_table=fric, restitution_table=rest)


def filter_tables(world: Any, raw: np.ndarray) -> Dict[str, Any]:
    """Dense filter slots and whether each slot pair collides (``model.py:143-155``)."""
    vals, slot = _dense_slots(raw)
    filts = [world.filter_for(int(v)) for v in vals]
    tab = np.array([[fa.collides_with(fb) for fb in filts] for fa in filts], np.uint8).reshape(len(vals), len(vals))
    return dict(filt_slot=slot, n_filt=len(vals), filter_table=tab)


def joint_arrays(constraints: List[Any]) -> Dict[str, Any]:
    """Body pairs, world anchors and given body-frame anchors of point joints
    (``joints.py:60-121``): bit 0 / 1 of ``joint_fixed`` marks ``local_a`` /
    ``local_b`` as set."""
    J = len(constraints)
    la, lb, fixed = np.zeros((J, 3)), np.zeros((J, 3)), np.zeros(J, np.uint8)
    for j, c in enumerate(constraints):
        for bit, loc, out in ((1, getattr(c, 'local_a', None), la), (2, getattr(c, 'local_b', None), lb)):
            if loc is not None:
                out[j] = np.asarray(loc, dtype='d').reshape(3)
                fixed[j] |= bit
    return dict(n_joints=J,
                joint_a=np.array([c.a for c in constraints], np.int32),
                joint_b=np.array([c.b for c in constraints], np.int32),
                joint_anchor=np.array([c.anchor for c in constraints], dtype='d').reshape(-1, 3),
                joint_local_a=la, joint_local_b=lb, joint_fixed=fixed)


def world_arrays(world: Any, velocity_iterations: Optional[int] = None,
                 with_joints: bool = True) -> Dict[str, Any]:
    """The world as flat contiguous arrays (row-major ``[n][3]`` / ``[n][4]``)."""
    n = int(world.body_count)
    f64 = lambda name: np.ascontiguousarray(getattr(world, name), dtype='d')
    i32 = lambda name: np.ascontiguousarray(getattr(world, name), dtype=np.int32)
    cs, ts = i32('collider_shape'), i32('trigger_shape')
    out: Dict[str, Any] = dict(n=n)
    for name in ('position', 'orientation', 'prev_position', 'prev_orientation', 'linear_velocity',
                 'angular_velocity', 'inv_mass', 'inv_inertia', 'gravity_factor', 'linear_damping',
                 'angular_damping', 'quadratic_drag', 'sleep_timer'):
        out[name] = f64(name)
    out.update(motion_type=i32('motion_type'), collider_shape=cs, trigger_shape=ts,
               awake=np.ascontiguousarray(world.awake, dtype=np.uint8))
    out.update(body_shapes(world.shapes, cs, ts))
    out.update(material_tables(world, i32('collider_material')))
    out.update(filter_tables(world, i32('collider_filter')))
    if velocity_iterations is None:
        velocity_iterations = int(getattr(world, 'velocity_iterations', 10))
    out.update(gravity=uniform_gravity(world),
               default_linear_damping=float(world.default_linear_damping),
               default_angular_damping=float(world.default_angular_damping),
               sleep_enabled=int(bool(world.sleep_enabled)), velocity_iterations=int(velocity_iterations),
               has_collision=int(world._collision is not None),
               joint_iterations=int(getattr(world, 'joint_iterations', 10)))
    out.update(joint_arrays(list(world.joint_constraints) if with_joints else []))
    return out


def layout_key(world: Any, velocity_iterations: int, with_joints: bool = True) -> Tuple:
    """What a device world's layout depends on: a change means reloading it."""
    mats = tuple((m.staticFriction, m.dynamicFriction, m.restitution, m.frictionCombine, m.restitutionCombine)
                 for m in world.materials)
    filts = tuple((tuple(f.collisionSystems), None if f.collideWithSystems is None else tuple(f.collideWithSystems),
                   tuple(f.notCollideWithSystems)) for f in world.filters)
    cols = tuple(np.ascontiguousarray(getattr(world, name), dtype=np.int32).tobytes()
                 for name in ('motion_type', 'collider_shape', 'trigger_shape', 'collider_material',
                              'collider_filter'))
    def local(x):
        return None if x is None else tuple(np.asarray(x, 'd').reshape(3).tolist())
    joints = tuple((c.a, c.b, tuple(np.asarray(c.anchor, 'd').tolist()), local(getattr(c, 'local_a', None)),
                    local(getattr(c, 'local_b', None))) for c in world.joint_constraints) if with_joints else ()
    return (int(world.body_count), len(world.shapes), mats, filts, tuple(sorted(world.pair_friction.items())),
            cols, joints, int(velocity_iterations), int(getattr(world, 'joint_iterations', 10)),
            world._collision is not None)
