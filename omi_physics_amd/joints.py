"""Point (ball) joints solved on the device after the contacts.

``PointConstraint(a, b, anchor, local_a=None, local_b=None)`` holds two bodies'
anchor points together (reference ``joints.py:60-121``); ``-1`` pins against
the world at ``anchor``.  A body-frame anchor left as ``None`` is taken from
the pose at the joint's first step and kept (as the reference's ``prepare``
stores it); a given one is used as is.  Each step the device wakes the joint
bodies, prepares every joint (effective-mass matrix K, Baumgarte bias 0.2) and
runs ``PhysicsWorld.joint_iterations`` Gauss-Seidel sweeps over the joints in
the order they were added (``world.py:599-611``), each a 3x3 solve
``K P = -(v_rel + bias/dt)``.
"""
from typing import Optional

import numpy as np


class PointConstraint:
    """Ball joint between bodies ``a`` and ``b`` (either may be ``-1``, the world)."""

    def __init__(self, a: int, b: int, anchor, local_a: Optional[np.ndarray] = None,
                 local_b: Optional[np.ndarray] = None) -> None:
        self.a, self.b = int(a), int(b)
        self.anchor = np.asarray(anchor, dtype='d')
        self.local_a = None if local_a is None else np.asarray(local_a, dtype='d').reshape(3)
        self.local_b = None if local_b is None else np.asarray(local_b, dtype='d').reshape(3)
