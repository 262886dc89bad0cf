"""OMI physics data model (shapes, motion, colliders, materials, filters, gravity).

Field names, defaults and semantics follow the reference's ``model.py`` so that a
scene built for ``omi_physics`` builds unchanged here:

* ``Shape`` / ``Motion`` / ``Collider`` / ``Trigger`` / ``Material`` /
  ``CollisionFilter`` / ``Gravity``  (reference ``model.py:32-178``);
* ``combine`` — the OMI friction/restitution combine rule (``model.py:231-243``);
* ``CollisionFilter.collides_with`` — symmetric layer/mask test (``model.py:143-155``).

Only what the contact hot path reads is carried; the joint/limit records of the
reference are not part of this package.
"""
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

# Motion types (OMI_physics_body.motion.type)
STATIC = 'static'
KINEMATIC = 'kinematic'
DYNAMIC = 'dynamic'

# Combine modes (OMI physics material)
AVERAGE = 'average'
MINIMUM = 'minimum'
MAXIMUM = 'maximum'
MULTIPLY = 'multiply'

#: combine mode -> integer code shared with the device material tables
COMBINE_CODE = {AVERAGE: 0, MINIMUM: 1, MAXIMUM: 2, MULTIPLY: 3}


@dataclass
class Shape:
    """One collision proxy, shared by index (reference ``model.py:32-82``)."""
    type: str = 'box'
    size: tuple = (1.0, 1.0, 1.0)               # box full extents
    radius: float = 0.5                          # sphere
    height: float = 2.0                          # capsule / cylinder
    radiusBottom: float = 0.5
    radiusTop: float = 0.5
    points: Optional[np.ndarray] = None          # convex vertices (M, 3)
    indices: Optional[np.ndarray] = None         # trimesh triangles (T, 3)
    mesh: int = -1

    @classmethod
    def box(cls, size: Sequence[float] = (1.0, 1.0, 1.0)) -> 'Shape':
        """A box collider with full extents ``size``."""
        return cls(type='box', size=tuple(float(s) for s in size))

    @classmethod
    def sphere(cls, radius: float = 0.5) -> 'Shape':
        """A sphere collider."""
        return cls(type='sphere', radius=float(radius))

    @classmethod
    def convex(cls, points) -> 'Shape':
        """A convex-hull collider from an ``(M, 3)`` local vertex cloud."""
        return cls(type='convex', points=np.asarray(points, dtype='d'))


@dataclass
class Motion:
    """Rigid-body dynamics parameters (reference ``model.py:85-102``)."""
    type: str = DYNAMIC
    mass: float = 1.0
    centerOfMass: tuple = (0.0, 0.0, 0.0)
    inertiaDiagonal: tuple = (0.0, 0.0, 0.0)     # 0 => derive from shape + mass
    inertiaOrientation: tuple = (0.0, 0.0, 0.0, 1.0)
    linearVelocity: tuple = (0.0, 0.0, 0.0)
    angularVelocity: tuple = (0.0, 0.0, 0.0)
    gravityFactor: float = 1.0
    linearDamping: float = 0.0
    angularDamping: float = 0.0
    quadraticDrag: float = 0.0


@dataclass
class Collider:
    """A solid shape that generates contacts (reference ``model.py:105-110``)."""
    shape: int = -1
    physicsMaterial: int = -1
    collisionFilter: int = -1


@dataclass
class Trigger:
    """A sensor shape: excluded from contact generation (reference ``model.py:113-117``)."""
    shape: int = -1
    collisionFilter: int = -1


@dataclass
class Material:
    """Friction / restitution with OMI combine modes (reference ``model.py:120-127``)."""
    staticFriction: float = 0.6
    dynamicFriction: float = 0.6
    restitution: float = 0.0
    frictionCombine: str = AVERAGE
    restitutionCombine: str = AVERAGE


@dataclass
class CollisionFilter:
    """Layer/mask filter (reference ``model.py:130-155``).

    ``collideWithSystems=None`` admits every system not explicitly excluded.
    """
    collisionSystems: tuple = ()
    collideWithSystems: Optional[tuple] = None
    notCollideWithSystems: tuple = ()

    def collides_with(self, other: 'CollisionFilter') -> bool:
        """True when both filters admit each other."""
        return _admits(self, other) and _admits(other, self)


def _admits(a: CollisionFilter, b: CollisionFilter) -> bool:
    """Whether ``a``'s policy admits a body that belongs to ``b``'s systems."""
    member = set(b.collisionSystems)
    if member.intersection(a.notCollideWithSystems):
        return False
    if a.collideWithSystems is None:
        return True
    return bool(member.intersection(a.collideWithSystems))


DIRECTIONAL = 'directional'


@dataclass
class Gravity:
    """Global directional gravity (reference ``model.py:163-178``)."""
    type: str = DIRECTIONAL
    gravity: float = 9.81
    direction: tuple = (0.0, -1.0, 0.0)
    priority: int = 0
    replace: bool = False
    stop: bool = False
    center: tuple = (0.0, 0.0, 0.0)


DEFAULT_MATERIAL = Material()
DEFAULT_FILTER = CollisionFilter()


def combine(a: float, b: float, mode: str) -> float:
    """OMI combine of two coefficients; unknown modes average (reference ``model.py:231-243``)."""
    if mode == MINIMUM:
        return min(a, b)
    if mode == MAXIMUM:
        return max(a, b)
    if mode == MULTIPLY:
        return a * b
    return 0.5 * (a + b)
