"""Compute-backend seam: the MI355X HIP backend (reference ``backend.py``).

``HipBackend`` offers the reference backend's surface — ``name``,
``integrate_forces(world, dt)``, ``integrate_positions(world, dt)``,
``refit_aabbs(world)`` (``backend.py:21-53``) — plus the collision-stage
object that ``PhysicsWorld`` installs as ``world._collision`` (the
``run(world, dt)`` duck type of ``world.py:642-676``).

Inside ``PhysicsWorld.step`` the seam calls only record which stages the step
asks for; the device work is one fused launch sequence (a HIP graph after the
first step) enqueued at ``end_step``, so the world's state never leaves HBM
between steps.  Outside a step, ``refit_aabbs`` launches the refit kernel on
its own.

``select_backend`` (``backend.py:95-120``) accepts ``'hip'``/``'gpu'``/
``'auto'``.  Unlike the reference there is no silent CPU fallback: the HIP
library must be present, and stepping without a visible GPU raises.
"""
from typing import Any, Optional

import numpy as np

from . import _ffi
from .batch import Batch, capacities_for, contact_list
from .triggers import TriggerSystem


class HipCollisionStages:
    """``world._collision``: broadphase -> narrowphase -> solver -> trigger overlaps, run on the
    device by the step; ``triggers`` keeps the overlap set across steps (reference
    ``world.py:640-676``)."""

    def __init__(self, backend: 'HipBackend') -> None:
        self.backend = backend
        self.triggers = TriggerSystem()

    def run(self, world: Any, dt: float) -> None:
        """Marks the collision stages for this step (they run in the fused device step)."""
        self.backend._stages.add('collision')


class HipBackend:
    """MI355X backend: the whole fixed step runs on one GPU, FP64 or mixed FP32."""

    name = 'hip'

    def __init__(self, precision: Optional[str] = None) -> None:
        _ffi.load()                     # fails loudly when the library is missing
        self.precision = precision
        self._batch: Optional[Batch] = None
        self._layout = -1
        self._stages = set()
        self._in_step = False
        self._has_triggers = False

    # -- device state ---------------------------------------------------------
    def _ensure(self, world: Any) -> Batch:
        """Create/refresh the one-world batch that holds ``world`` in HBM."""
        prec = self.precision or getattr(world, 'precision', 'fp64')
        if self._batch is None or self._layout != world._layout_version or self._batch.precision != prec:
            arrays = world.export_arrays()
            n = int(arrays['n'])
            caps = capacities_for(max(n, 1))
            hull = max(1, len(np.asarray(arrays['hull_pts']).reshape(-1, 3)))
            if self._batch is not None:
                self._batch.close()
            self._batch = Batch(1, max(n, 1), precision=prec, max_hull_pts=hull,
                                max_joints=max(1, int(arrays['n_joints'])), record_broadphase=True, **caps)
            if n:
                self._batch.load(0, arrays)
            self._has_triggers = bool(n and (np.asarray(arrays['trigger_shape']) >= 0).any())
            self._layout = world._layout_version
            world._host_dirty = False
        elif world._host_dirty and world._n:
            n = world._n
            self._batch.set_state(0, world._position[:n], world._orientation[:n],
                                  world._linear_velocity[:n], world._angular_velocity[:n],
                                  world._awake[:n], world._sleep_timer[:n])
            world._host_dirty = False
        return self._batch

    @property
    def batch(self) -> Optional[Batch]:
        return self._batch

    def make_collision_stages(self, world: Any) -> HipCollisionStages:
        return HipCollisionStages(self)

    # -- the reference seam -----------------------------------------------------
    def begin_step(self, world: Any, dt: float) -> None:
        self._ensure(world)
        self._stages = set()
        self._in_step = True

    def integrate_forces(self, world: Any, dt: float) -> None:
        """Gravity, damping, drag for awake dynamic bodies (reference ``backend.py:26-40``)."""
        self._stages.add('forces')

    def refit_aabbs(self, world: Any, margin: float = 0.05) -> None:
        """World AABBs from pose (reference ``world.py:471-508``)."""
        if self._in_step:
            self._stages.add('refit')
            return
        b = self._ensure(world)
        if world._n:
            b.refit(margin)

    def integrate_positions(self, world: Any, dt: float) -> None:
        """Symplectic position/orientation update (reference ``backend.py:42-49``)."""
        self._stages.add('positions')

    def solve_joints(self, world: Any, dt: float) -> None:
        """Point joints after the contacts (reference ``world.py:599-611``)."""
        self._stages.add('joints')

    def update_sleep(self, world: Any, dt: float) -> None:
        """Island-aware sleeping (reference ``world.py:510-567``)."""
        self._stages.add('sleep')

    def end_step(self, world: Any, dt: float) -> None:
        self._in_step = False
        if world._n:
            self._batch.step(dt, 1)
            col = world._collision
            if col is not None and (self._has_triggers or col.triggers.overlaps):
                # trigger events need this step's overlap set: the one wait of a
                # step with sensors (reference world.py:674-676)
                self._batch.sync()
                for event in col.triggers.update(map(tuple, self._batch.trigger_overlaps(0))):
                    world.dispatch_trigger(*event)

    # -- host mirror --------------------------------------------------------------
    def download(self, world: Any) -> None:
        """Copy the device state into the world's host columns."""
        if self._batch is None or not world._n:
            return
        n = world._n
        self._batch.sync()
        s = self._batch.read(0)
        world._position[:n] = s['position']
        world._orientation[:n] = s['orientation']
        world._prev_position[:n] = s['prev_position']
        world._prev_orientation[:n] = s['prev_orientation']
        world._linear_velocity[:n] = s['linear_velocity']
        world._angular_velocity[:n] = s['angular_velocity']
        world._aabb_min[:n] = s['aabb_min']
        world._aabb_max[:n] = s['aabb_max']
        world._awake[:n] = s['awake'].astype(bool)
        world._sleep_timer[:n] = s['sleep_timer']

    def read_contacts(self, world: Any):
        if self._batch is None or not world._n:
            return []
        self._batch.sync()
        return contact_list(self._batch.contacts(0))

    def sync(self) -> None:
        if self._batch is not None:
            self._batch.sync()


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
