"""``PhysicsWorld`` — structure-of-arrays world state and the fixed-step pipeline.

The public surface mirrors the reference ``world.py``: the same constructor
keywords (``world.py:30-35``), ``add_shape/add_material/add_filter``
(``:175-193``), ``set_pair_friction`` (``:207-223``), ``add_body``
(``:253-299``) with the same mass-property rules (``:301-316``, ``:614-639``),
the same column views (``:145-167``) and the same ``step(dt)`` stage order
(``:410-430``).

The difference is where the columns live while the world steps.  The compute
backend (``backend.HipBackend``) keeps a device-resident copy of every column in
HBM and runs each stage there; the host numpy columns are a mirror that is
refreshed lazily when a column view is read, and re-uploaded before the next
step when it may have been edited.  ``step()`` itself never copies state
between host and device.
"""
from typing import Any, Dict, List, Optional, Tuple
import os

import numpy as np

from . import model
from . import mathutil

_TYPE_INT = {model.STATIC: 0, model.KINEMATIC: 1, model.DYNAMIC: 2}
STATIC, KINEMATIC, DYNAMIC = 0, 1, 2

#: shape-kind codes shared with the device and the C ABI
KIND_NONE, KIND_SPHERE, KIND_BOX, KIND_CONVEX = 0, 1, 2, 3

# Constants of the contact pipeline (reference values).
REFIT_MARGIN = 0.05                 # world.py:471
SLEEP_LIN, SLEEP_ANG, SLEEP_TIME = 0.08, 0.25, 0.4   # world.py:510-511
WAKE_SPEED = 0.15                   # solver.py:322


class PhysicsWorld:
    """A rigid-body world: add shapes/materials/filters, then bodies, then step."""

    _VEC = ('position', 'prev_position', 'orientation', 'prev_orientation',
            'linear_velocity', 'angular_velocity', 'inv_inertia',
            'aabb_min', 'aabb_max')
    _SCALAR_F = ('inv_mass', 'mass', 'gravity_factor', 'linear_damping',
                 'angular_damping', 'quadratic_drag', 'sleep_timer',
                 'restitution_pad')
    _SCALAR_I = ('motion_type', 'collider_shape', 'collider_material',
                 'collider_filter', 'trigger_shape', 'island')
    #: columns the device pipeline writes; the host mirror refreshes these
    _DEVICE_WRITES = ('position', 'prev_position', 'orientation', 'prev_orientation',
                      'linear_velocity', 'angular_velocity', 'aabb_min', 'aabb_max',
                      'sleep_timer', 'awake')

    def __init__(self, gravity: Optional[model.Gravity] = None,
                 fixed_dt: float = 1.0 / 60.0, max_frame: float = 0.25,
                 backend: Any = None, sleep_enabled: bool = True,
                 default_linear_damping: float = 0.0,
                 default_angular_damping: float = 0.0,
                 gpu_threshold: int = 10000,
                 velocity_iterations: int = 10,
                 precision: str = 'fp64') -> None:
        self.gravity = gravity if gravity is not None else model.Gravity()
        self.fixed_dt = fixed_dt
        self.max_frame = max_frame
        self.sleep_enabled = sleep_enabled
        self.default_linear_damping = default_linear_damping
        self.default_angular_damping = default_angular_damping
        self.gpu_threshold = gpu_threshold
        self.velocity_iterations = int(velocity_iterations)
        self.precision = precision

        self.shapes: List[model.Shape] = []
        self.materials: List[model.Material] = []
        self.filters: List[model.CollisionFilter] = []
        self.pair_friction: Dict[Tuple[int, int], Tuple[float, float]] = {}
        self.bodies: List[Any] = []

        self._n = 0
        self._cap = 0
        self._accumulator = 0.0
        self.time = 0.0
        self._alloc(8)
        self._collision: Any = None
        self.trigger_listeners: List[Any] = []
        self.joint_constraints: List[Any] = []
        self.joint_iterations = 10
        self._layout_version = 0           # bumps when bodies/tables change
        self._host_dirty = True            # host mirror may be newer than device
        self._device_fresh = False         # device state newer than host mirror
        self._contacts_cache: Any = None
        self._init_backend(backend)

    # -- storage ---------------------------------------------------------
    def _alloc(self, cap: int) -> None:
        """Allocate zeroed columns for ``cap`` bodies (reference ``world.py:110-122``)."""
        self._cap = cap
        for name in self._VEC:
            width = 4 if 'orientation' in name else 3
            setattr(self, '_' + name, np.zeros((cap, width), dtype='d'))
        for name in self._SCALAR_F:
            setattr(self, '_' + name, np.zeros(cap, dtype='d'))
        for name in self._SCALAR_I:
            setattr(self, '_' + name, np.full(cap, -1, dtype='i4'))
        self._awake = np.ones(cap, dtype=bool)
        self._orientation[:, 3] = 1.0
        self._prev_orientation[:, 3] = 1.0

    def _grow(self) -> None:
        """Double the capacity, preserving state (reference ``world.py:124-138``)."""
        old = self._cap
        new = max(8, old * 2)
        for name in self._VEC + self._SCALAR_F + self._SCALAR_I + ('awake',):
            arr = getattr(self, '_' + name)
            grown = np.zeros((new,) + arr.shape[1:], dtype=arr.dtype)
            if name in self._SCALAR_I:
                grown.fill(-1)
            grown[:old] = arr
            setattr(self, '_' + name, grown)
        self._orientation[old:, 3] = 1.0
        self._prev_orientation[old:, 3] = 1.0
        self._awake[old:] = True
        self._cap = new

    def _pull(self) -> None:
        """Refresh the host mirror from the device if the device is newer."""
        if self._device_fresh:
            self._device_fresh = False
            self.backend.download(self)
        # a view handed out may be written through: upload before the next step
        self._host_dirty = True

    def _view(self, name: str) -> np.ndarray:
        """Column ``name`` over the live bodies (the first ``body_count`` rows)."""
        if name in self._DEVICE_WRITES:
            self._pull()
        return getattr(self, '_' + name)[:self._n]

    position = property(lambda s: s._view('position'))
    prev_position = property(lambda s: s._view('prev_position'))
    orientation = property(lambda s: s._view('orientation'))
    prev_orientation = property(lambda s: s._view('prev_orientation'))
    linear_velocity = property(lambda s: s._view('linear_velocity'))
    angular_velocity = property(lambda s: s._view('angular_velocity'))
    inv_mass = property(lambda s: s._view('inv_mass'))
    mass = property(lambda s: s._view('mass'))
    inv_inertia = property(lambda s: s._view('inv_inertia'))
    gravity_factor = property(lambda s: s._view('gravity_factor'))
    linear_damping = property(lambda s: s._view('linear_damping'))
    angular_damping = property(lambda s: s._view('angular_damping'))
    quadratic_drag = property(lambda s: s._view('quadratic_drag'))
    motion_type = property(lambda s: s._view('motion_type'))
    collider_shape = property(lambda s: s._view('collider_shape'))
    collider_material = property(lambda s: s._view('collider_material'))
    collider_filter = property(lambda s: s._view('collider_filter'))
    trigger_shape = property(lambda s: s._view('trigger_shape'))
    aabb_min = property(lambda s: s._view('aabb_min'))
    aabb_max = property(lambda s: s._view('aabb_max'))
    awake = property(lambda s: s._view('awake'))
    island = property(lambda s: s._view('island'))
    sleep_timer = property(lambda s: s._view('sleep_timer'))

    @property
    def body_count(self) -> int:
        """The number of live bodies."""
        return self._n

    # -- document tables -------------------------------------------------
    def add_shape(self, shape: model.Shape) -> int:
        """Register a collision shape; return its index."""
        if shape.type not in ('box', 'sphere', 'convex'):
            raise ValueError('shape type %r is not supported by the contact pipeline '
                             '(box, sphere and convex are)' % shape.type)
        self.shapes.append(shape)
        return len(self.shapes) - 1

    def add_material(self, mat: model.Material) -> int:
        """Register a physics material; return its index."""
        self.materials.append(mat)
        self._layout_version += 1
        return len(self.materials) - 1

    def add_filter(self, filt: model.CollisionFilter) -> int:
        """Register a collision filter; return its index."""
        self.filters.append(filt)
        self._layout_version += 1
        return len(self.filters) - 1

    def material_for(self, index: int) -> model.Material:
        """The material at ``index``, or the default material if out of range."""
        if 0 <= index < len(self.materials):
            return self.materials[index]
        return model.DEFAULT_MATERIAL

    def filter_for(self, index: int) -> model.CollisionFilter:
        """The filter at ``index``, or the default filter if out of range."""
        if 0 <= index < len(self.filters):
            return self.filters[index]
        return model.DEFAULT_FILTER

    def set_pair_friction(self, material_a: int, material_b: int,
                          static_mu: float, dynamic_mu: float) -> None:
        """Override contact friction for one unordered material pair (reference ``world.py:207-218``)."""
        key = (material_a, material_b) if material_a <= material_b else (material_b, material_a)
        self.pair_friction[key] = (float(static_mu), float(dynamic_mu))
        self._layout_version += 1

    def pair_friction_for(self, material_a: int,
                          material_b: int) -> Optional[Tuple[float, float]]:
        """``(static_mu, dynamic_mu)`` for the pair, or ``None``."""
        key = (material_a, material_b) if material_a <= material_b else (material_b, material_a)
        return self.pair_friction.get(key)

    # -- bodies ----------------------------------------------------------
    def add_body(self, motion: Optional[model.Motion] = None,
                 collider: Optional[model.Collider] = None,
                 trigger: Optional[model.Trigger] = None,
                 position=(0, 0, 0), orientation=(0, 0, 0, 1),
                 handle: Any = None) -> int:
        """Create a body from its motion/collider/trigger; return its index."""
        self._pull()
        if motion is None:
            motion = model.Motion()
        if self._n >= self._cap:
            self._grow()
        i = self._n
        self._n += 1
        self._position[i] = position
        self._prev_position[i] = position
        self._orientation[i] = orientation
        self._prev_orientation[i] = orientation
        self._linear_velocity[i] = motion.linearVelocity
        self._angular_velocity[i] = motion.angularVelocity
        self._motion_type[i] = _TYPE_INT[motion.type]
        self._gravity_factor[i] = motion.gravityFactor
        self._linear_damping[i] = motion.linearDamping
        self._angular_damping[i] = motion.angularDamping
        self._quadratic_drag[i] = motion.quadraticDrag
        self._awake[i] = True
        self._sleep_timer[i] = 0.0

        shape: Optional[model.Shape] = None
        if collider is not None and 0 <= collider.shape < len(self.shapes):
            shape = self.shapes[collider.shape]
            self._collider_shape[i] = collider.shape
            self._collider_material[i] = collider.physicsMaterial
            self._collider_filter[i] = collider.collisionFilter
            self._collision = self._collision or self.backend.make_collision_stages(self)
        if trigger is not None and 0 <= trigger.shape < len(self.shapes):
            self._trigger_shape[i] = trigger.shape
            self._collision = self._collision or self.backend.make_collision_stages(self)
        self._set_mass_properties(i, motion, shape)
        self.bodies.append(handle)
        self._layout_version += 1
        return i

    def _set_mass_properties(self, i: int, motion: model.Motion,
                             shape: Optional[model.Shape]) -> None:
        """Mass and inverse inertia; zero for static/kinematic (reference ``world.py:301-316``)."""
        if motion.type != model.DYNAMIC:
            self._inv_mass[i] = 0.0
            self._mass[i] = 0.0
            self._inv_inertia[i] = 0.0
            return
        m = max(motion.mass, 1e-9)
        self._mass[i] = m
        self._inv_mass[i] = 1.0 / m
        diag = np.asarray(motion.inertiaDiagonal, dtype='d')
        if not diag.any():
            diag = inertia_diagonal(shape, m) if shape is not None else np.zeros(3)
        with np.errstate(divide='ignore'):
            self._inv_inertia[i] = np.where(diag > 0, 1.0 / np.maximum(diag, 1e-12), 0.0)

    # -- backend policy --------------------------------------------------
    def _init_backend(self, backend: Any) -> None:
        """Choose the compute backend (reference ``world.py:319-350``).

        An explicit backend object is used as-is.  Otherwise the name (or the
        ``OPENGLCONTEXT_PHYSICS_BACKEND`` variable) goes to
        :func:`backend.select_backend`, which raises when the HIP library
        cannot be loaded rather than substituting a CPU path.
        """
        from .backend import select_backend
        if backend is not None and not isinstance(backend, str):
            self.backend = backend
        else:
            mode = backend or os.environ.get('OPENGLCONTEXT_PHYSICS_BACKEND', 'auto')
            self.backend = select_backend(mode)
        self._backend_mode = 'fixed'

    # -- masks & gravity -------------------------------------------------
    def dynamic_mask(self) -> np.ndarray:
        """Awake dynamic bodies."""
        return (self.motion_type == DYNAMIC) & self.awake

    def moving_mask(self) -> np.ndarray:
        """Awake dynamic or kinematic bodies."""
        return ((self.motion_type == DYNAMIC) | (self.motion_type == KINEMATIC)) & self.awake

    def resolve_gravity(self) -> np.ndarray:
        """The uniform gravity acceleration vector ``(3,)`` (reference ``world.py:386-396``)."""
        base = np.asarray(self.gravity.direction, dtype='d')
        norm = np.linalg.norm(base)
        base = base / norm if norm else base
        return base * self.gravity.gravity

    def wake(self, i: int) -> None:
        """Wake body ``i`` and reset its sleep timer."""
        self._pull()
        self._awake[i] = True
        self._sleep_timer[i] = 0.0

    # -- triggers --------------------------------------------------------
    def add_trigger_listener(self, fn: Any) -> None:
        """Register ``fn(event_type, trigger_index, other_index)`` (reference ``world.py:580-582``)."""
        self.trigger_listeners.append(fn)

    def dispatch_trigger(self, event_type: str, trigger_index: int, other_index: int) -> None:
        """Notify every registered trigger listener of one overlap event (reference ``world.py:584-588``)."""
        for fn in self.trigger_listeners:
            fn(event_type, trigger_index, other_index)

    def add_joint_constraint(self, constraint: Any) -> Any:
        """Register a ``joints.PointConstraint``, waking its bodies (reference ``world.py:590-597``)."""
        from .joints import PointConstraint
        if not isinstance(constraint, PointConstraint):
            raise TypeError('only PointConstraint joints run on the device')
        self.joint_constraints.append(constraint)
        for i in (constraint.a, constraint.b):
            if i >= 0:
                self.wake(i)
        self._layout_version += 1
        return constraint

    # -- the step --------------------------------------------------------
    def advance(self, real_dt: float) -> float:
        """Fixed-timestep accumulator (reference ``world.py:399-408``)."""
        self._accumulator = min(self._accumulator + real_dt, self.max_frame)
        while self._accumulator >= self.fixed_dt:
            self.step(self.fixed_dt)
            self._accumulator -= self.fixed_dt
        return self._accumulator / self.fixed_dt

    def step(self, dt: float) -> None:
        """Advance one fixed step: forces, collision, positions, sleep (reference ``world.py:410-430``).

        Every stage is a device launch on the backend's stream; nothing here
        waits for the GPU.
        """
        self.backend.begin_step(self, dt)
        self.backend.integrate_forces(self, dt)
        if self._collision is not None:
            self.backend.refit_aabbs(self)
            self._collision.run(self, dt)
        self.backend.integrate_positions(self, dt)
        if self.joint_constraints:
            self.backend.solve_joints(self, dt)
        if self.sleep_enabled:
            self.backend.update_sleep(self, dt)
        self.backend.end_step(self, dt)
        self._device_fresh = True
        self._contacts_cache = None
        self.time += dt

    def refit_aabbs(self, margin: float = REFIT_MARGIN) -> None:
        """Recompute world AABBs on the device (reference ``world.py:471-508``)."""
        self.backend.refit_aabbs(self, margin)
        self._device_fresh = True

    @property
    def contacts(self) -> List[Any]:
        """The last step's contacts, materialised on demand (reference ``world.py:672``)."""
        if self._contacts_cache is None:
            self._contacts_cache = self.backend.read_contacts(self) if self._collision else []
        return self._contacts_cache

    def writeback(self, alpha: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
        """Interpolated render poses (reference ``world.py:432-437``)."""
        pos = self.prev_position * (1 - alpha) + self.position * alpha
        quat = mathutil.quat_normalize(
            self.prev_orientation * (1 - alpha) + self.orientation * alpha)
        return pos, quat

    # -- export for the device loader -------------------------------------
    def export_arrays(self) -> Dict[str, Any]:
        """The world as flat contiguous arrays in the layout the C ABI loads.

        Also used by the tests to hand the identical state to the oracle.
        """
        return export_world_arrays(self)


def inertia_diagonal(shape: Optional[model.Shape], mass: float) -> np.ndarray:
    """Solid-body inertia diagonal about the centre (reference ``world.py:614-639``)."""
    if shape is None:
        return np.zeros(3)
    t = shape.type
    if t == 'sphere':
        v = 0.4 * mass * shape.radius ** 2
        return np.array([v, v, v])
    if t == 'box':
        x, y, z = shape.size
        k = mass / 12.0
        return k * np.array([y * y + z * z, x * x + z * z, x * x + y * y])
    if shape.points is not None and len(shape.points):
        lo = shape.points.min(axis=0)
        hi = shape.points.max(axis=0)
        x, y, z = hi - lo
        k = mass / 12.0
        return k * np.array([y * y + z * z, x * x + z * z, x * x + y * y])
    return np.zeros(3)


def _slot_table(raw: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Distinct raw index values -> dense slots; returns ``(values, slot_per_body)``."""
    values, slots = np.unique(raw, return_inverse=True)
    return values.astype(np.int64), slots.astype(np.int32)


def export_world_arrays(w: PhysicsWorld) -> Dict[str, Any]:
    """Flatten a world into the C-ABI layout (row-major ``[n][3]`` / ``[n][4]`` columns).

    Material and filter indices are remapped to dense slots over the distinct
    raw values the bodies use, and the per-slot-pair friction, restitution
    (reference ``solver.py:252-259``) and filter (``broadphase.py:297-301``)
    outcomes are tabulated, so device and oracle read one table entry per pair.
    """
    n = w._n
    g = lambda name: np.ascontiguousarray(getattr(w, '_' + name)[:n])
    kind = np.zeros(n, dtype=np.int32)
    half = np.zeros((n, 3), dtype='d')
    hull_off = np.full(n, -1, dtype=np.int32)
    hull_cnt = np.zeros(n, dtype=np.int32)
    hull_pts: List[np.ndarray] = []
    hull_total = 0
    cs = w._collider_shape[:n]
    ts = w._trigger_shape[:n]
    for i in range(n):
        idx = int(cs[i]) if cs[i] >= 0 else int(ts[i])
        if idx < 0:
            continue
        if cs[i] >= 0 and ts[i] >= 0 and cs[i] != ts[i]:
            # one shape per body on the device: its collider's, which is also
            # what the trigger overlap test uses
            raise NotImplementedError('body %d has a collider and a different trigger shape' % i)
        sh = w.shapes[idx]
        if sh.type == 'sphere':
            kind[i] = KIND_SPHERE
            half[i] = sh.radius
        elif sh.type == 'box':
            kind[i] = KIND_BOX
            half[i] = 0.5 * np.asarray(sh.size, dtype='d')
        else:
            kind[i] = KIND_CONVEX
            pts = np.ascontiguousarray(sh.points, dtype='d')
            hull_off[i] = hull_total
            hull_cnt[i] = len(pts)
            hull_pts.append(pts)
            hull_total += len(pts)
    mat_vals, mat_slot = _slot_table(w._collider_material[:n])
    nm = len(mat_vals)
    fric = np.zeros((nm, nm), dtype='d')
    rest = np.zeros((nm, nm), dtype='d')
    for i, ra in enumerate(mat_vals):
        ma = w.material_for(int(ra))
        for j, rb in enumerate(mat_vals):
            mb = w.material_for(int(rb))
            rest[i, j] = model.combine(ma.restitution, mb.restitution, ma.restitutionCombine)
            override = w.pair_friction_for(int(ra), int(rb))
            fric[i, j] = (override[1] if override is not None else
                          model.combine(ma.dynamicFriction, mb.dynamicFriction,
                                        ma.frictionCombine))
    filt_vals, filt_slot = _slot_table(w._collider_filter[:n])
    nf = len(filt_vals)
    ftab = np.zeros((nf, nf), dtype=np.uint8)
    for i, ra in enumerate(filt_vals):
        for j, rb in enumerate(filt_vals):
            ftab[i, j] = w.filter_for(int(ra)).collides_with(w.filter_for(int(rb)))
    return dict(
        n=n,
        position=g('position'), orientation=g('orientation'),
        prev_position=g('prev_position'), prev_orientation=g('prev_orientation'),
        linear_velocity=g('linear_velocity'), angular_velocity=g('angular_velocity'),
        inv_mass=g('inv_mass'), inv_inertia=g('inv_inertia'),
        gravity_factor=g('gravity_factor'), linear_damping=g('linear_damping'),
        angular_damping=g('angular_damping'), quadratic_drag=g('quadratic_drag'),
        sleep_timer=g('sleep_timer'),
        motion_type=g('motion_type').astype(np.int32),
        collider_shape=g('collider_shape').astype(np.int32),
        trigger_shape=g('trigger_shape').astype(np.int32),
        awake=np.ascontiguousarray(w._awake[:n]).astype(np.uint8),
        kind=kind, half=half,
        hull_off=hull_off, hull_cnt=hull_cnt,
        hull_pts=(np.concatenate(hull_pts) if hull_pts else np.zeros((0, 3))),
        mat_slot=mat_slot, n_mat=nm,
        friction_table=np.ascontiguousarray(fric), restitution_table=np.ascontiguousarray(rest),
        filt_slot=filt_slot, n_filt=nf, filter_table=np.ascontiguousarray(ftab),
        gravity=np.asarray(w.resolve_gravity(), dtype='d'),
        default_linear_damping=float(w.default_linear_damping),
        default_angular_damping=float(w.default_angular_damping),
        sleep_enabled=int(bool(w.sleep_enabled)),
        velocity_iterations=int(w.velocity_iterations),
        has_collision=int(w._collision is not None),
        n_joints=len(w.joint_constraints), joint_iterations=int(w.joint_iterations),
        joint_a=np.array([c.a for c in w.joint_constraints], np.int32),
        joint_b=np.array([c.b for c in w.joint_constraints], np.int32),
        joint_anchor=np.array([c.anchor for c in w.joint_constraints], dtype='d').reshape(-1, 3),
    )
