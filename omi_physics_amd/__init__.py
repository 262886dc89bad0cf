"""omi_physics_amd — the omi_physics contact hot path, native to MI355X (gfx950).

Same user surface as the reference package for the contact pipeline
(``PhysicsWorld``, ``model``, ``mathutil``, ``select_backend``), with every
stage of the step — AABB refit, sweep-and-prune, SAT box/sphere contacts,
the sequential-impulse solver, integration and sleeping — running as
hand-written HIP kernels on world state resident in HBM.  ``Batch`` steps many
independent worlds in one launch sequence; ``serve`` admits worlds
continuously.
"""
from . import model
from . import mathutil
from .world import PhysicsWorld
from .backend import HipBackend, select_backend
from .batch import Batch, Contact, batch_from_arrays

__version__ = "0.1.0"

__all__ = ["PhysicsWorld", "HipBackend", "select_backend", "Batch", "Contact",
           "batch_from_arrays", "model", "mathutil", "__version__"]
