"""Compute-backend seam: the MI355X HIP backend (reference ``backend.py:21-53``).

``HipBackend`` runs worlds on the GPU two ways.

*Stage path* — any world whose NumPy columns are authoritative between the
calls of its own ``step()``, the reference's ``PhysicsWorld`` included.
``HipBackend`` is the reference backend duck type: ``name``,
``integrate_forces(world, dt)``, ``refit_aabbs(world)`` and
``integrate_positions(world, dt)`` each run that stage's kernel on the world's
columns (``omi_integrate_forces``, ``omi_refit``, ``omi_integrate_positions``)
and write them back.  The first of these calls on a world installs a
:class:`HipCollisionStages` as ``world._collision`` in place of the reference's
``_CollisionStages`` (keeping its trigger overlaps and its solver's iteration
count).  Its ``run(world, dt)`` — the reference's ``world.py:657-676`` —
uploads the state, runs broadphase, solid filter, narrowphase, wake, contact
solve and trigger overlaps on a one-world device batch (``omi_batch_collide``)
whose warm-start cache stays on the device across steps and world edits, writes
positions, velocities and wake flags back, sets ``world.contacts`` and
dispatches trigger events before the host world goes on to its joints,
``integrate_positions`` and ``_update_sleep``.

*Resident path* — ``omi_physics_amd.PhysicsWorld`` keeps its state in HBM and
steps with one fused, graph-captured device step (``step_resident``).  Host
edits of its columns are found by comparing them with the values last
synchronised and are uploaded (state, parameters) or reloaded (layout, with the
warm-start cache and the joints' body-frame anchors kept) before the next step.

``select_backend`` (``backend.py:95-120``) accepts ``'hip'``/``'gpu'``/
``'auto'``.  There is no CPU fallback: the HIP library must load, and stepping
without a visible GPU raises.
"""
from typing import Any, Dict, Optional, Tuple

import numpy as np

from . import _ffi, layout
from .batch import Batch, capacities_for, contact_list
from .triggers import TriggerSystem

STATE_COLUMNS = ('position', 'orientation', 'linear_velocity', 'angular_velocity', 'awake', 'sleep_timer')
PARAM_COLUMNS = ('inv_mass', 'inv_inertia', 'gravity_factor', 'linear_damping', 'angular_damping',
                 'quadratic_drag')


def _pow2(n: int, lo: int = 64) -> int:
    p = lo
    while p < n:
        p <<= 1
    return p


class DeviceWorld:
    """One world held in slot 0 of its own device batch, reloaded on layout changes
    with its warm-start cache (and joint anchors) carried over."""

    def __init__(self, precision: str, with_joints: bool) -> None:
        self.precision = precision
        self.with_joints = with_joints
        self.batch: Optional[Batch] = None
        self.key: Optional[Tuple] = None
        self.n = 0
        self.n_joints = 0
        self.has_triggers = False
        self.grow = 1                       # capacity multiplier after a capacity error
        self.error: Optional[str] = None    # set after a capacity error of a resident step
        self.snapshot: Dict[str, np.ndarray] = {}
        self.settings: Optional[Tuple] = None

    def close(self) -> None:
        if self.batch is not None:
            self.batch.close()
            self.batch = None
        self.key = None

    def load(self, arrays: Dict[str, Any], key: Tuple, keep_cache: bool = True) -> None:
        n = int(arrays['n'])
        npts = len(np.asarray(arrays['hull_pts']).reshape(-1, 3))
        nj = int(arrays['n_joints'])
        cache = None
        if keep_cache and self.batch is not None and self.key is not None and self.n:
            cache = self.batch.read_cache(0)
            keep = (cache['a'] < n) & (cache['b'] < n)
            cache = {k: v[keep] for k, v in cache.items()}
        cfg = self.batch.config if self.batch is not None else None
        need = (_pow2(n) * self.grow, _pow2(max(npts, 1), 16), _pow2(max(nj, 1), 16))
        if cfg is None or cfg.max_bodies < need[0] or cfg.max_hull_pts < need[1] or cfg.max_joints < need[2]:
            if self.batch is not None:
                self.batch.close()
            caps = capacities_for(need[0])
            caps = {k: v * self.grow for k, v in caps.items()}
            self.batch = Batch(1, need[0], precision=self.precision, max_hull_pts=need[1], max_joints=need[2],
                               record_broadphase=True, **caps)
        self.batch.load(0, arrays)
        if cache is not None and len(cache['a']):
            self.batch.write_cache(0, cache)
        self.key = key
        self.n = n
        self.n_joints = nj
        self.has_triggers = bool(n and (np.asarray(arrays['trigger_shape']) >= 0).any())
        self.error = None


def _settings(world: Any) -> Tuple:
    g = world.gravity
    return (tuple(np.asarray(g.direction, 'd').tolist()), float(g.gravity), float(world.default_linear_damping),
            float(world.default_angular_damping), bool(world.sleep_enabled), int(world.velocity_iterations))


class HipCollisionStages:
    """``world._collision`` on the GPU for a world with host-authoritative columns:
    broadphase -> narrowphase -> wake -> contact solve -> trigger events
    (reference ``world.py:640-676``)."""

    def __init__(self, backend: 'HipBackend', triggers: Optional[TriggerSystem] = None,
                 velocity_iterations: int = 10) -> None:
        self.backend = backend
        self.triggers = triggers if triggers is not None else TriggerSystem()
        self.velocity_iterations = int(velocity_iterations)
        self.contacts: list = []
        self.device: Optional[DeviceWorld] = None

    def _ready(self, world: Any) -> DeviceWorld:
        if self.device is None:
            prec = self.backend.precision or getattr(world, 'precision', 'fp64')
            self.device = DeviceWorld(prec, with_joints=False)
        dev = self.device
        key = layout.layout_key(world, self.velocity_iterations, with_joints=False)
        if dev.key != key:
            dev.load(layout.world_arrays(world, velocity_iterations=self.velocity_iterations, with_joints=False), key)
        return dev

    def run(self, world: Any, dt: float) -> None:
        """One collision pass on the device; results land in the world's columns."""
        n = int(world.body_count)
        self.contacts = []
        if n == 0:
            world.contacts = self.contacts
            return
        dev = self._ready(world)
        for attempt in range(4):
            b = dev.batch
            b.set_state(0, world.position, world.orientation, world.linear_velocity, world.angular_velocity,
                        world.awake, world.sleep_timer)
            b.set_aabbs(0, world.aabb_min, world.aabb_max)
            b.set_params(0, **{name: getattr(world, name) for name in PARAM_COLUMNS})
            b.collide(dt)
            try:
                b.sync()
                break
            except RuntimeError as e:
                if 'capacity' not in str(e) or attempt == 3:
                    raise
                # the host columns still hold the pre-collision state and the failed step
                # kept the previous cache (a world that overflows gets no contacts): rerun
                # with twice the room
                dev.grow *= 2
                dev.load(layout.world_arrays(world, velocity_iterations=self.velocity_iterations, with_joints=False),
                         dev.key)
        st = b.read(0)
        world.position[:] = st['position']
        world.linear_velocity[:] = st['linear_velocity']
        world.angular_velocity[:] = st['angular_velocity']
        world.awake[:] = st['awake'].astype(bool)
        world.sleep_timer[:] = st['sleep_timer']
        self.contacts = contact_list(b.contacts(0))
        world.contacts = self.contacts
        if dev.has_triggers or self.triggers.overlaps:
            for event in self.triggers.update(map(tuple, b.trigger_overlaps(0))):
                world.dispatch_trigger(*event)


class HipBackend:
    """The MI355X backend: stage path for host-column worlds, resident path for
    ``omi_physics_amd.PhysicsWorld``; FP64 or mixed FP32."""

    name = 'hip'

    def __init__(self, precision: Optional[str] = None) -> None:
        _ffi.load()                     # fails loudly when the library is missing
        self.precision = precision

    # -- the reference seam (stage path) ----------------------------------------
    def _adopt(self, world: Any) -> None:
        """Install the device collision stage in place of a host one."""
        col = getattr(world, '_collision', None)
        if col is not None and not isinstance(col, HipCollisionStages):
            trig = TriggerSystem()
            trig.overlaps = set(getattr(getattr(col, 'triggers', None), 'overlaps', set()))
            iters = getattr(getattr(col, 'solver', None), 'velocity_iterations', 10)
            world._collision = HipCollisionStages(self, trig, iters)

    def integrate_forces(self, world: Any, dt: float) -> None:
        """Gravity, damping and quadratic drag of awake dynamic bodies (reference ``backend.py:26-40``)."""
        self._adopt(world)
        n = int(world.body_count)
        if n == 0:
            return
        L = _ffi.load()
        lv = _ffi.c64(world.linear_velocity)
        av = _ffi.c64(world.angular_velocity)
        g = _ffi.c64(np.broadcast_to(np.asarray(world.resolve_gravity(), dtype='d'), (n, 3)))
        mt = _ffi.ci32(world.motion_type)
        aw = np.ascontiguousarray(world.awake, dtype=np.uint8)
        cols = [_ffi.c64(getattr(world, k)) for k in ('gravity_factor', 'linear_damping', 'angular_damping',
                                                      'quadratic_drag')]
        _ffi.check(L.omi_integrate_forces(n, _ffi.ptr(mt, _ffi._i32), _ffi.ptr(aw, _ffi._u8), _ffi.ptr(g, _ffi._d),
                                          *[_ffi.ptr(c, _ffi._d) for c in cols],
                                          float(world.default_linear_damping), float(world.default_angular_damping),
                                          float(dt), _ffi.ptr(lv, _ffi._d), _ffi.ptr(av, _ffi._d)))
        world.linear_velocity[:] = lv
        world.angular_velocity[:] = av

    def refit_aabbs(self, world: Any, margin: float = 0.05) -> None:
        """World AABBs from pose and shape (reference ``world.py:471-508``)."""
        self._adopt(world)
        n = int(world.body_count)
        if n == 0:
            return
        L = _ffi.load()
        sh = layout.body_shapes(world.shapes, np.asarray(world.collider_shape), np.asarray(world.trigger_shape))
        pos, quat = _ffi.c64(world.position), _ffi.c64(world.orientation)
        lo, hi = np.empty((n, 3)), np.empty((n, 3))
        pts = _ffi.c64(sh['hull_pts'].reshape(-1, 3) if len(sh['hull_pts']) else np.zeros((1, 3)))
        _ffi.check(L.omi_refit(n, _ffi.ptr(pos, _ffi._d), _ffi.ptr(quat, _ffi._d), _ffi.ptr(sh['kind'], _ffi._i32),
                               _ffi.ptr(_ffi.c64(sh['half']), _ffi._d), _ffi.ptr(sh['hull_off'], _ffi._i32),
                               _ffi.ptr(sh['hull_cnt'], _ffi._i32), _ffi.ptr(pts, _ffi._d), len(sh['hull_pts']),
                               float(margin), _ffi.ptr(lo, _ffi._d), _ffi.ptr(hi, _ffi._d)))
        world.aabb_min[:] = lo
        world.aabb_max[:] = hi

    def integrate_positions(self, world: Any, dt: float) -> None:
        """x += v dt and the quaternion update of moving bodies (reference ``backend.py:42-49``)."""
        n = int(world.body_count)
        if n == 0:
            return
        L = _ffi.load()
        pos, quat = _ffi.c64(world.position), _ffi.c64(world.orientation)
        lv, av = _ffi.c64(world.linear_velocity), _ffi.c64(world.angular_velocity)
        mt = _ffi.ci32(world.motion_type)
        aw = np.ascontiguousarray(world.awake, dtype=np.uint8)
        _ffi.check(L.omi_integrate_positions(n, _ffi.ptr(mt, _ffi._i32), _ffi.ptr(aw, _ffi._u8),
                                             _ffi.ptr(lv, _ffi._d), _ffi.ptr(av, _ffi._d), float(dt),
                                             _ffi.ptr(pos, _ffi._d), _ffi.ptr(quat, _ffi._d)))
        world.position[:] = pos
        world.orientation[:] = quat

    def make_collision_stages(self, world: Any) -> HipCollisionStages:
        return HipCollisionStages(self, velocity_iterations=int(getattr(world, 'velocity_iterations', 10)))

    # -- resident path -------------------------------------------------------------
    def _device(self, world: Any) -> DeviceWorld:
        dev = world.__dict__.get('_hip_device')
        if dev is None:
            prec = self.precision or getattr(world, 'precision', 'fp64')
            dev = DeviceWorld(prec, with_joints=True)
            world._hip_device = dev
        return dev

    def _keep_joint_anchors(self, world: Any, dev: DeviceWorld) -> None:
        """Store the body-frame anchors the device fixed at the joints' first step
        on the constraint objects (as the reference's ``prepare`` does), so a
        reload keeps them."""
        if dev.batch is None or not dev.n_joints:
            return
        la, lb, fx = dev.batch.read_joint_anchors(0, dev.n_joints)
        for j, c in enumerate(world.joint_constraints[:dev.n_joints]):
            if fx[j] & 1 and getattr(c, 'local_a', None) is None:
                c.local_a = la[j].copy()
            if fx[j] & 2 and getattr(c, 'local_b', None) is None:
                c.local_b = lb[j].copy()

    def _upload_edits(self, world: Any) -> DeviceWorld:
        """Bring the device copy up to date with the host edits since the last sync."""
        dev = self._device(world)
        if dev.error is not None:
            raise RuntimeError(dev.error)
        edited = world._edited
        if world.body_count == 0:
            edited.clear()
            return dev
        key = dev.key
        if dev.batch is None or 'layout' in edited:
            key = layout.layout_key(world, world.velocity_iterations)
        if dev.batch is None or key != dev.key:
            self._keep_joint_anchors(world, dev)
            dev.load(layout.world_arrays(world), layout.layout_key(world, world.velocity_iterations))
            dev.snapshot = {k: np.array(world._raw(k)) for k in STATE_COLUMNS + PARAM_COLUMNS}
            dev.settings = _settings(world)
        elif dev.n:
            b = dev.batch
            if 'state' in edited:
                cur = {k: world._raw(k) for k in STATE_COLUMNS}
                if any(not np.array_equal(cur[k], dev.snapshot[k]) for k in STATE_COLUMNS):
                    b.set_state(0, *[cur[k] for k in STATE_COLUMNS])
                    dev.snapshot.update({k: np.array(v) for k, v in cur.items()})
            settings = _settings(world)
            params_changed = False
            if 'params' in edited:
                cur = {k: world._raw(k) for k in PARAM_COLUMNS}
                params_changed = any(not np.array_equal(cur[k], dev.snapshot[k]) for k in PARAM_COLUMNS)
            if params_changed or settings != dev.settings:
                kw = {k: world._raw(k) for k in PARAM_COLUMNS}
                b.set_params(0, gravity=layout.uniform_gravity(world),
                             default_damping=(world.default_linear_damping, world.default_angular_damping),
                             sleep_enabled=world.sleep_enabled, velocity_iterations=world.velocity_iterations, **kw)
                dev.snapshot.update({k: np.array(v) for k, v in kw.items()})
                dev.settings = settings
        edited.clear()
        return dev

    def step_resident(self, world: Any, dt: float) -> None:
        """Enqueue one fused step of a resident world (reference stage order,
        ``world.py:410-430``, one captured graph); with sensors, wait for it and
        dispatch the trigger events."""
        dev = self._upload_edits(world)
        if not dev.n:
            return
        dev.batch.step(dt, 1)
        col = world._collision
        if col is not None and (dev.has_triggers or col.triggers.overlaps):
            self._sync(dev)
            for event in col.triggers.update(map(tuple, dev.batch.trigger_overlaps(0))):
                world.dispatch_trigger(*event)

    def refit_resident(self, world: Any, margin: float) -> None:
        dev = self._upload_edits(world)
        if dev.n:
            dev.batch.refit(margin)

    def _sync(self, dev: DeviceWorld) -> None:
        try:
            dev.batch.sync()
        except RuntimeError as e:
            if 'capacity' in str(e):
                dev.error = ('a step ran out of device pair/contact capacity, so the world state on the device is '
                             'invalid (%s); call world.recover_device() to reload it from the host mirror into a '
                             'larger batch' % e)
                raise RuntimeError(dev.error) from None
            raise

    def download(self, world: Any) -> None:
        """Copy the device state into the world's host mirror (the mirror is left
        untouched if the device state is invalid)."""
        dev = self._device(world)
        if dev.error is not None:
            raise RuntimeError(dev.error)
        if dev.batch is None or not dev.n:
            return
        self._sync(dev)
        st = dev.batch.read(0)
        for k in ('position', 'orientation', 'prev_position', 'prev_orientation', 'linear_velocity',
                  'angular_velocity', 'aabb_min', 'aabb_max', 'sleep_timer'):
            world._raw(k)[:] = st[k]
        world._raw('awake')[:] = st['awake'].astype(bool)
        dev.snapshot.update({k: np.array(world._raw(k)) for k in STATE_COLUMNS})

    def read_contacts(self, world: Any):
        dev = self._device(world)
        if dev.batch is None or not dev.n:
            return []
        self._sync(dev)
        return contact_list(dev.batch.contacts(0))

    def recover(self, world: Any) -> None:
        """Forget the device copy after a capacity error; the next step reloads the
        host mirror into a batch with doubled capacities."""
        dev = self._device(world)
        grow = dev.grow * 2
        dev.close()
        dev.grow = grow
        dev.error = None
        dev.n = 0

    def world_batch(self, world: Any) -> Optional[Batch]:
        """The device batch holding a resident world (slot 0), once it has stepped."""
        return self._device(world).batch


def select_backend(prefer: str = 'auto') -> HipBackend:
    """The compute backend for ``prefer`` (reference ``backend.py:95-120``).

    ``'hip'``, ``'gpu'``, ``'auto'`` (and the reference's ``'glcompute'``) give
    :class:`HipBackend`; ``'hip-fp32'`` gives the mixed-precision form.  CPU
    names raise: the CPU path of this project is the test oracle, not a
    backend.
    """
    if prefer in ('hip', 'gpu', 'auto', 'glcompute', None):
        return HipBackend()
    if prefer == 'hip-fp32':
        return HipBackend(precision='fp32')
    raise RuntimeError('backend %r is not available: this package runs the contact pipeline on the '
                       'GPU only (use "hip")' % (prefer,))
