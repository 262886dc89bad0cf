"""Point (ball) joints solved on the device after the contacts.

``PointConstraint(a, b, anchor)`` holds two bodies' anchor points together
(reference ``joints.py:60-121``); ``-1`` pins against the world at
``anchor``.  Each step the device wakes the joint bodies, prepares every joint
(local anchors fixed at the first step, effective-mass matrix K, Baumgarte
bias 0.2) and runs ``PhysicsWorld.joint_iterations`` Gauss-Seidel sweeps over
the joints in the order they were added (``world.py:599-611``), each a 3x3
solve ``K P = -(ċ + bias/dt)``.
"""
from typing import Optional

import numpy as np


class PointConstraint:
    """Ball joint between bodies ``a`` and ``b`` (either may be ``-1``, the world)."""

    def __init__(self, a: int, b: int, anchor, local_a: Optional[np.ndarray] = None,
                 local_b: Optional[np.ndarray] = None) -> None:
        if local_a is not None or local_b is not None:
            raise NotImplementedError('explicit local anchors are not supported; they are taken '
                                      'from the pose at the first step, like the reference default')
        self.a, self.b = int(a), int(b)
        self.anchor = np.asarray(anchor, dtype='d')
