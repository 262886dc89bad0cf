"""``PhysicsWorld`` with its body state resident in HBM.

The API is the reference ``PhysicsWorld``'s (``world.py:30-35`` constructor
keywords, ``add_shape/add_material/add_filter``, ``set_pair_friction``,
``add_body`` and its mass rules, the column views, ``step``/``advance``/
``writeback``, trigger listeners and point joints), so scenes written for the
reference build here unchanged.

What differs is where the columns live.  The device copy, held by the
``HipBackend``, is the live state: ``step()`` enqueues one fused device step
(every stage of the reference's ``step``, ``world.py:410-430``, in its order)
and returns without waiting.  The host keeps a mirror, one numpy array per
column in a :class:`BodyColumns` store: a view of a column the device writes
first refreshes the mirror, and handing out a view marks its column group
(state, parameters or layout) as possibly edited.  Before the next step the
backend compares each marked group with the values it last synchronised and
uploads what changed — or, for layout, reloads the world with its warm-start
cache and joint anchors kept.
"""
from typing import Any, Callable, Dict, List, Optional, Sequence, Tuple
import os

import numpy as np

from . import layout, model
from . import mathutil
from .layout import KIND_NONE, KIND_SPHERE, KIND_BOX, KIND_CONVEX  # noqa: F401  (re-exported codes)

STATIC, KINEMATIC, DYNAMIC = 0, 1, 2
_MOTION_CODE = {model.STATIC: STATIC, model.KINEMATIC: KINEMATIC, model.DYNAMIC: DYNAMIC}

# Constants of the contact pipeline (reference values).
REFIT_MARGIN = 0.05                 # world.py:471
SLEEP_LIN, SLEEP_ANG, SLEEP_TIME = 0.08, 0.25, 0.4   # world.py:510-511
WAKE_SPEED = 0.15                   # solver.py:322


class BodyColumns:
    """Per-body columns, row ``i`` = body ``i``, in arrays whose row capacity is a
    power of two; new rows start at each column's rest value.

    ``groups`` tells which kind of device upload an edit of a column needs:
    ``'state'`` (pose, velocity, sleep), ``'params'`` (mass, inertia, forces)
    or ``'layout'`` (what the world's device tables are built from).
    """

    #: name -> (row shape, dtype, rest value, group)
    SPEC: Dict[str, Tuple[Tuple[int, ...], Any, Any, str]] = {
        'position': ((3,), 'f8', 0.0, 'state'),
        'prev_position': ((3,), 'f8', 0.0, 'state'),
        'orientation': ((4,), 'f8', (0.0, 0.0, 0.0, 1.0), 'state'),
        'prev_orientation': ((4,), 'f8', (0.0, 0.0, 0.0, 1.0), 'state'),
        'linear_velocity': ((3,), 'f8', 0.0, 'state'),
        'angular_velocity': ((3,), 'f8', 0.0, 'state'),
        'awake': ((), bool, True, 'state'),
        'sleep_timer': ((), 'f8', 0.0, 'state'),
        'aabb_min': ((3,), 'f8', 0.0, 'state'),
        'aabb_max': ((3,), 'f8', 0.0, 'state'),
        'inv_mass': ((), 'f8', 0.0, 'params'),
        'mass': ((), 'f8', 0.0, 'params'),
        'inv_inertia': ((3,), 'f8', 0.0, 'params'),
        'gravity_factor': ((), 'f8', 0.0, 'params'),
        'linear_damping': ((), 'f8', 0.0, 'params'),
        'angular_damping': ((), 'f8', 0.0, 'params'),
        'quadratic_drag': ((), 'f8', 0.0, 'params'),
        'restitution_pad': ((), 'f8', 0.0, 'params'),
        'motion_type': ((), 'i4', -1, 'layout'),
        'collider_shape': ((), 'i4', -1, 'layout'),
        'collider_material': ((), 'i4', -1, 'layout'),
        'collider_filter': ((), 'i4', -1, 'layout'),
        'trigger_shape': ((), 'i4', -1, 'layout'),
        'island': ((), 'i4', -1, 'params'),
    }
    #: columns the device step writes
    DEVICE_WRITES = ('position', 'prev_position', 'orientation', 'prev_orientation', 'linear_velocity',
                     'angular_velocity', 'awake', 'sleep_timer', 'aabb_min', 'aabb_max')

    def __init__(self) -> None:
        self.n = 0
        self.rows = 0
        self.arrays: Dict[str, np.ndarray] = {}
        self._resize(8)

    def _resize(self, rows: int) -> None:
        for name, (shape, dtype, rest, _) in self.SPEC.items():
            fresh = np.empty((rows,) + shape, dtype=dtype)
            fresh[...] = rest
            old = self.arrays.get(name)
            if old is not None:
                fresh[:self.n] = old[:self.n]
            self.arrays[name] = fresh
        self.rows = rows

    def append(self) -> int:
        """A new row at its rest values; returns its index."""
        if self.n == self.rows:
            self._resize(2 * self.rows)
        self.n += 1
        return self.n - 1

    def live(self, name: str) -> np.ndarray:
        return self.arrays[name][:self.n]


def inertia_diagonal(shape: Optional[model.Shape], mass: float) -> np.ndarray:
    """Solid-body inertia about the centre for a shape of ``mass`` (reference
    ``world.py:614-639``): sphere 2/5 m r^2; box m/12 (y^2+z^2, x^2+z^2, x^2+y^2)
    of its full extents; a convex shape takes the box of its points' bounds."""
    def box(extent):
        x, y, z = extent
        return (mass / 12.0) * np.array([y * y + z * z, x * x + z * z, x * x + y * y])
    if shape is None:
        return np.zeros(3)
    if shape.type == 'sphere':
        return np.full(3, 0.4 * mass * shape.radius ** 2)
    if shape.type == 'box':
        return box(shape.size)
    if shape.points is not None and len(shape.points):
        return box(shape.points.max(axis=0) - shape.points.min(axis=0))
    return np.zeros(3)


def mass_properties(motion: model.Motion, shape: Optional[model.Shape]) -> Tuple[float, float, np.ndarray]:
    """``(mass, inv_mass, inv_inertia)`` of a body (reference ``world.py:301-316``):
    static and kinematic bodies get zeros; a dynamic body's mass is at least 1e-9,
    its inertia diagonal the given one or the shape's, inverted where positive."""
    if motion.type != model.DYNAMIC:
        return 0.0, 0.0, np.zeros(3)
    m = max(motion.mass, 1e-9)
    diag = np.asarray(motion.inertiaDiagonal, dtype='d')
    if not diag.any():
        diag = inertia_diagonal(shape, m)
    with np.errstate(divide='ignore'):
        inv_i = np.where(diag > 0, 1.0 / np.maximum(diag, 1e-12), 0.0)
    return m, 1.0 / m, inv_i


def _column_view(name: str):
    return property(lambda self: self._view(name), doc='Column %r over the live bodies.' % name)


class PhysicsWorld:
    """A rigid-body world on the GPU: add shapes/materials/filters, then bodies, then step."""

    def __init__(self, gravity: Optional[model.Gravity] = None,
                 fixed_dt: float = 1.0 / 60.0, max_frame: float = 0.25,
                 backend: Any = None, sleep_enabled: bool = True,
                 default_linear_damping: float = 0.0,
                 default_angular_damping: float = 0.0,
                 gpu_threshold: int = 10000,
                 velocity_iterations: int = 10,
                 precision: str = 'fp64') -> None:
        self.gravity = gravity if gravity is not None else model.Gravity()
        self.fixed_dt, self.max_frame = fixed_dt, max_frame
        self.sleep_enabled = sleep_enabled
        self.default_linear_damping = default_linear_damping
        self.default_angular_damping = default_angular_damping
        self.gpu_threshold = gpu_threshold          # accepted for the reference's signature
        self.velocity_iterations = int(velocity_iterations)
        self.precision = precision
        self.shapes: List[model.Shape] = []
        self.materials: List[model.Material] = []
        self.filters: List[model.CollisionFilter] = []
        self.pair_friction: Dict[Tuple[int, int], Tuple[float, float]] = {}
        self.bodies: List[Any] = []
        self.trigger_listeners: List[Callable[..., None]] = []
        self.joint_constraints: List[Any] = []
        self.joint_iterations = 10
        self.time = 0.0
        self._accumulator = 0.0
        self._cols = BodyColumns()
        self._collision: Any = None
        # device synchronisation: column groups edited on the host since the last
        # upload, and whether the device has stepped past the host mirror
        self._edited = {'state', 'params', 'layout'}
        self._device_fresh = False
        self._contacts_cache: Optional[List[Any]] = None
        self.backend = self._make_backend(backend)
        self._backend_mode = 'fixed'

    @staticmethod
    def _make_backend(backend: Any) -> Any:
        """The compute backend: an object is used as is; a name (or
        ``OPENGLCONTEXT_PHYSICS_BACKEND``, reference ``world.py:319-350``) goes to
        :func:`backend.select_backend`, which fails loudly without the HIP library."""
        if backend is not None and not isinstance(backend, str):
            return backend
        from .backend import select_backend
        return select_backend(backend or os.environ.get('OPENGLCONTEXT_PHYSICS_BACKEND', 'auto'))

    # -- columns ---------------------------------------------------------
    def _view(self, name: str) -> np.ndarray:
        if name in BodyColumns.DEVICE_WRITES:
            self._refresh_mirror()
        self._edited.add(BodyColumns.SPEC[name][3])     # a view may be written through
        return self._cols.live(name)

    def _refresh_mirror(self) -> None:
        if self._device_fresh:
            self.backend.download(self)                 # raises if the device state is invalid
            self._device_fresh = False

    def recover_device(self) -> None:
        """After a capacity error: drop the device copy; the next step reloads the
        world from the host mirror (the state last read back) into a batch with
        doubled pair and contact capacities."""
        self.backend.recover(self)
        self._device_fresh = False
        self._edited.update(('state', 'params', 'layout'))

    def _raw(self, name: str) -> np.ndarray:
        """The host mirror of a column without refreshing it or marking an edit."""
        return self._cols.live(name)

    @property
    def body_count(self) -> int:
        """The number of live bodies."""
        return self._cols.n

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
        self._edited.add('layout')
        return len(self.materials) - 1

    def add_filter(self, filt: model.CollisionFilter) -> int:
        """Register a collision filter; return its index."""
        self.filters.append(filt)
        self._edited.add('layout')
        return len(self.filters) - 1

    def material_for(self, index: int) -> model.Material:
        """Material ``index``; an index outside the table means the default material."""
        return self.materials[index] if 0 <= index < len(self.materials) else model.DEFAULT_MATERIAL

    def filter_for(self, index: int) -> model.CollisionFilter:
        """Filter ``index``; an index outside the table means the default filter."""
        return self.filters[index] if 0 <= index < len(self.filters) else model.DEFAULT_FILTER

    def set_pair_friction(self, material_a: int, material_b: int,
                          static_mu: float, dynamic_mu: float) -> None:
        """Friction for contacts between two materials, either order (reference ``world.py:207-218``)."""
        self.pair_friction[(min(material_a, material_b), max(material_a, material_b))] = \
            (float(static_mu), float(dynamic_mu))
        self._edited.add('layout')

    def pair_friction_for(self, material_a: int, material_b: int) -> Optional[Tuple[float, float]]:
        """``(static_mu, dynamic_mu)`` set for the material pair, or ``None``."""
        return self.pair_friction.get((min(material_a, material_b), max(material_a, material_b)))

    # -- bodies ----------------------------------------------------------
    def add_body(self, motion: Optional[model.Motion] = None,
                 collider: Optional[model.Collider] = None,
                 trigger: Optional[model.Trigger] = None,
                 position=(0, 0, 0), orientation=(0, 0, 0, 1),
                 handle: Any = None) -> int:
        """Append a body built from its motion, collider and trigger; return its index."""
        self._refresh_mirror()
        motion = motion or model.Motion()
        i = self._cols.append()
        col = self._cols.arrays
        for name in ('position', 'prev_position'):
            col[name][i] = position
        for name in ('orientation', 'prev_orientation'):
            col[name][i] = orientation
        col['linear_velocity'][i] = motion.linearVelocity
        col['angular_velocity'][i] = motion.angularVelocity
        col['motion_type'][i] = _MOTION_CODE[motion.type]
        col['gravity_factor'][i] = motion.gravityFactor
        col['linear_damping'][i] = motion.linearDamping
        col['angular_damping'][i] = motion.angularDamping
        col['quadratic_drag'][i] = motion.quadraticDrag
        shape = None
        if collider is not None and 0 <= collider.shape < len(self.shapes):
            shape = self.shapes[collider.shape]
            col['collider_shape'][i] = collider.shape
            col['collider_material'][i] = collider.physicsMaterial
            col['collider_filter'][i] = collider.collisionFilter
        if trigger is not None and 0 <= trigger.shape < len(self.shapes):
            col['trigger_shape'][i] = trigger.shape
        if col['collider_shape'][i] >= 0 or col['trigger_shape'][i] >= 0:
            self._collision = self._collision or self.backend.make_collision_stages(self)
        col['mass'][i], col['inv_mass'][i], col['inv_inertia'][i] = mass_properties(motion, shape)
        self.bodies.append(handle)
        self._edited.update(('state', 'params', 'layout'))
        return i

    def apply_impulse(self, i: int, impulse) -> None:
        """``v += J / m`` on body ``i`` and wake it (reference ``world.py:231-238``)."""
        lv = self.linear_velocity
        lv[i] = lv[i] + self._raw('inv_mass')[i] * np.asarray(impulse, dtype='d')
        self.wake(i)

    def apply_angular_impulse(self, i: int, impulse) -> None:
        """``w += I_world^-1 L`` on body ``i`` about world axes, and wake it
        (reference ``world.py:240-251``)."""
        R = mathutil.quat_to_matrix(self.orientation[i])
        inv_world = (R * self._raw('inv_inertia')[i]) @ R.T
        av = self.angular_velocity
        av[i] = av[i] + inv_world @ np.asarray(impulse, dtype='d')
        self.wake(i)

    def wake(self, i: int) -> None:
        """Wake body ``i`` and reset its sleep timer."""
        self.awake[i] = True
        self.sleep_timer[i] = 0.0

    # -- masks & gravity -------------------------------------------------
    def dynamic_mask(self) -> np.ndarray:
        """The awake dynamic bodies."""
        return (self.motion_type == DYNAMIC) & self.awake

    def moving_mask(self) -> np.ndarray:
        """The awake dynamic or kinematic bodies."""
        return np.isin(self.motion_type, (DYNAMIC, KINEMATIC)) & self.awake

    def is_awake_mover(self, i: int) -> bool:
        """An awake dynamic body, or any kinematic body (reference ``world.py:381-384``)."""
        t = self._raw('motion_type')[i]
        return bool(t == KINEMATIC or (t == DYNAMIC and self.awake[i]))

    def resolve_gravity(self) -> np.ndarray:
        """The gravity acceleration ``(3,)`` every body feels (no gravity volumes here)."""
        return layout.uniform_gravity(self)

    # -- triggers and joints ---------------------------------------------
    def add_trigger_listener(self, fn: Callable[..., None]) -> None:
        """Register ``fn(event_type, trigger_index, other_index)``."""
        self.trigger_listeners.append(fn)

    def dispatch_trigger(self, event_type: str, trigger_index: int, other_index: int) -> None:
        """Call every trigger listener with one overlap event."""
        for fn in self.trigger_listeners:
            fn(event_type, trigger_index, other_index)

    def add_joint_constraint(self, constraint: Any) -> Any:
        """Register a ``joints.PointConstraint`` and wake its bodies."""
        from .joints import PointConstraint
        if not isinstance(constraint, PointConstraint):
            raise TypeError('the device solves PointConstraint joints')
        self.joint_constraints.append(constraint)
        for i in (constraint.a, constraint.b):
            if i >= 0:
                self.wake(i)
        self._edited.add('layout')
        return constraint

    # -- stepping --------------------------------------------------------
    def advance(self, real_dt: float) -> float:
        """Step as many fixed steps as the accumulated real time allows (clamped to
        ``max_frame``); return the leftover fraction of a step for render blending."""
        self._accumulator = min(self._accumulator + real_dt, self.max_frame)
        while self._accumulator >= self.fixed_dt:
            self.step(self.fixed_dt)
            self._accumulator -= self.fixed_dt
        return self._accumulator / self.fixed_dt

    def step(self, dt: float) -> None:
        """One fixed step of every stage, enqueued on the device without waiting."""
        self.backend.step_resident(self, dt)
        self._device_fresh = True
        self._contacts_cache = None
        self.time += dt

    def refit_aabbs(self, margin: float = REFIT_MARGIN) -> None:
        """Recompute the world AABBs from the current poses."""
        self.backend.refit_resident(self, margin)
        self._device_fresh = True

    @property
    def contacts(self) -> List[Any]:
        """The last step's contacts in solver order, read from the device on first use."""
        if self._contacts_cache is None:
            self._contacts_cache = self.backend.read_contacts(self) if self._collision is not None else []
        return self._contacts_cache

    @contacts.setter
    def contacts(self, value: List[Any]) -> None:
        self._contacts_cache = list(value)

    def writeback(self, alpha: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
        """Render poses blended from the previous to the current step by ``alpha``
        (positions linearly, orientations normalised after blending)."""
        a = float(alpha)
        pos = self.prev_position * (1.0 - a) + self.position * a
        quat = mathutil.quat_normalize(self.prev_orientation * (1.0 - a) + self.orientation * a)
        return pos, quat

    # -- export ----------------------------------------------------------
    def export_arrays(self) -> Dict[str, Any]:
        """The world in the C-ABI load layout (also what the oracle loads)."""
        return layout.world_arrays(self)


for _name in BodyColumns.SPEC:
    setattr(PhysicsWorld, _name, _column_view(_name))
del _name


def export_world_arrays(world: Any) -> Dict[str, Any]:
    """The C-ABI load layout of any world with the reference's public surface."""
    return layout.world_arrays(world)
