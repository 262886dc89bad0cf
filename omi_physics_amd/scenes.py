"""Seeded scene generators for the benchmark and parity workloads.

A scene is first produced as a plain *spec* (shapes, materials, bodies, world
keywords) and then built into a world by :func:`build`.  The builder only uses
the OMI data-model vocabulary (``Shape.box``, ``Motion(type=...)``,
``Collider(shape=, physicsMaterial=)``, ``add_body(...)``), so the same spec
builds an ``omi_physics_amd.PhysicsWorld`` or any world class with the
reference's interface — which is how the golden vectors and the parity tests
get identical scenes.

Workloads (sizes as named in each function):
  * ``boxes_on_ground``   — 1000 unit boxes dropped on a 20x20 m static ground.
  * ``pyramids_and_debris`` — 20 ten-level box pyramids plus 8000 boxes/spheres.
  * ``batched_worlds``    — 4096 worlds with lognormal body counts (median 64).
  * ``serving_requests``  — Poisson arrivals of worlds of 8..4096 bodies.

Randomness is ``numpy.random.default_rng(seed)`` throughout.
"""
from typing import Any, Dict, List, Optional

import numpy as np

GROUND_HALF = (10.0, 0.5, 10.0)


def _axis_angle_quats(rng: np.random.Generator, n: int) -> np.ndarray:
    """Orientations from a uniform axis on S² and a uniform angle in [0, 2π)."""
    axis = rng.normal(size=(n, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    ang = rng.uniform(0.0, 2.0 * np.pi, size=n)
    s = np.sin(0.5 * ang)
    return np.concatenate([axis * s[:, None], np.cos(0.5 * ang)[:, None]], axis=1)


def _ground(spec: Dict[str, Any], half=GROUND_HALF, position=(0.0, 0.0, 0.0), material=0) -> None:
    spec['shapes'].append(('box', tuple(2.0 * h for h in half)))
    spec['bodies'].append(dict(type='static', shape=len(spec['shapes']) - 1, mass=0.0,
                               position=tuple(position), orientation=(0.0, 0.0, 0.0, 1.0),
                               material=material))


def new_spec(friction: float = 0.5, restitution: float = 0.0, **world_kwargs) -> Dict[str, Any]:
    """An empty spec with one material and the given world keywords."""
    return dict(shapes=[], bodies=[],
                materials=[dict(staticFriction=friction, dynamicFriction=friction,
                                restitution=restitution)],
                world=dict(world_kwargs))


def boxes_on_ground(n: int = 1000, seed: int = 1, spread: float = 10.0,
                    y_range=(2.0, 50.0), sphere_fraction: float = 0.0,
                    sphere_radius: float = 0.5) -> Dict[str, Any]:
    """Unit boxes (and optionally spheres), mass 1, over a static (10, 0.5, 10) ground.

    With the defaults this is the single-world FP64 workload: 1000 boxes at
    x, z ~ U[-10, 10], y ~ U[2, 50], axis-angle orientations, μ = 0.5, e = 0.
    """
    rng = np.random.default_rng(seed)
    spec = new_spec()
    _ground(spec)
    spec['shapes'].append(('box', (1.0, 1.0, 1.0)))
    box_shape = len(spec['shapes']) - 1
    spec['shapes'].append(('sphere', float(sphere_radius)))
    sph_shape = len(spec['shapes']) - 1
    xz = rng.uniform(-spread, spread, size=(n, 2))
    y = rng.uniform(y_range[0], y_range[1], size=n)
    q = _axis_angle_quats(rng, n)
    is_sphere = rng.uniform(size=n) < sphere_fraction
    for i in range(n):
        spec['bodies'].append(dict(type='dynamic', shape=sph_shape if is_sphere[i] else box_shape,
                                   mass=1.0, position=(xz[i, 0], y[i], xz[i, 1]),
                                   orientation=tuple(q[i]), material=0))
    return spec


def pyramids_and_debris(n_pyramids: int = 20, levels: int = 10, n_debris: int = 8000,
                        seed: int = 2, footprint: float = 100.0) -> Dict[str, Any]:
    """Box pyramids (levels 10, 9, ..., 1 in one plane) plus box/sphere debris.

    The ground is one static box covering the footprint (half extents
    (55, 0.5, 55)); pyramids stand in a row along x spaced 4 m apart past
    their base width; debris is half boxes (half extents U[0.25, 1]) and
    half spheres (r ~ U[0.25, 1]), density 1, over the footprint at
    y ~ U[1, 80].
    """
    rng = np.random.default_rng(seed)
    spec = new_spec()
    half_fp = 0.5 * footprint + 5.0
    _ground(spec, half=(half_fp, 0.5, half_fp))
    spec['shapes'].append(('box', (1.0, 1.0, 1.0)))
    unit = len(spec['shapes']) - 1
    pitch = levels + 4.0
    x0 = -0.5 * pitch * (n_pyramids - 1)
    for p in range(n_pyramids):
        cx = x0 + p * pitch
        for lv in range(levels):
            count = levels - lv
            for k in range(count):
                x = cx + (k - 0.5 * (count - 1)) * 1.0
                spec['bodies'].append(dict(type='dynamic', shape=unit, mass=1.0,
                                           position=(x, 0.5 + lv * 1.0, -40.0),
                                           orientation=(0.0, 0.0, 0.0, 1.0), material=0))
    xz = rng.uniform(-0.5 * footprint, 0.5 * footprint, size=(n_debris, 2))
    y = rng.uniform(1.0, 80.0, size=n_debris)
    q = _axis_angle_quats(rng, n_debris)
    for i in range(n_debris):
        if i % 2 == 0:
            h = rng.uniform(0.25, 1.0, size=3)
            spec['shapes'].append(('box', tuple(2.0 * h)))
            mass = float(8.0 * h[0] * h[1] * h[2])
        else:
            r = float(rng.uniform(0.25, 1.0))
            spec['shapes'].append(('sphere', r))
            mass = float(4.0 / 3.0 * np.pi * r ** 3)
        spec['bodies'].append(dict(type='dynamic', shape=len(spec['shapes']) - 1, mass=mass,
                                   position=(xz[i, 0], y[i], xz[i, 1]),
                                   orientation=tuple(q[i]), material=0))
    return spec


def convex_heap(n: int = 2000, seed: int = 4, spread: float = 10.0, y_range=(1.0, 30.0),
                pts_range=(16, 32), axes_range=(0.3, 1.0)) -> Dict[str, Any]:
    """Convex hulls over a static ground: each hull is 16-32 points on an
    ellipsoid with semi-axes U[0.3, 1.0], cooked with the reference's hull
    algorithm (``hull.convex_hull``), mass 1, dropped at y ~ U[1, 30]."""
    from .hull import cook_convex
    rng = np.random.default_rng(seed)
    spec = new_spec()
    _ground(spec)
    xz = rng.uniform(-spread, spread, size=(n, 2))
    y = rng.uniform(y_range[0], y_range[1], size=n)
    q = _axis_angle_quats(rng, n)
    for i in range(n):
        m = int(rng.integers(pts_range[0], pts_range[1] + 1))
        axes = rng.uniform(axes_range[0], axes_range[1], size=3)
        d = rng.normal(size=(m, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        spec['shapes'].append(('convex', cook_convex(d * axes)))
        spec['bodies'].append(dict(type='dynamic', shape=len(spec['shapes']) - 1, mass=1.0,
                                   position=(xz[i, 0], y[i], xz[i, 1]), orientation=tuple(q[i]), material=0))
    return spec


def world_sizes(n_worlds: int = 4096, median: int = 64, sigma: float = 1.0,
                lo: int = 8, hi: int = 2048, seed: int = 3) -> np.ndarray:
    """Dynamic-body count per world: clip(round(lognormal(ln median, sigma)), lo, hi)."""
    rng = np.random.default_rng(seed)
    n = np.rint(rng.lognormal(np.log(median), sigma, size=n_worlds)).astype(np.int64)
    return np.clip(n, lo, hi)


def batched_worlds(n_worlds: int = 4096, seed: int = 3, sigma: float = 1.0,
                   sizes: Optional[np.ndarray] = None) -> List[Dict[str, Any]]:
    """Independent worlds, each a ground box plus a 50/50 box/sphere mix (unit
    boxes, r = 0.5 spheres, mass 1) placed like ``boxes_on_ground``; world ``w``
    is seeded ``seed * 100003 + w``."""
    if sizes is None:
        sizes = world_sizes(n_worlds, sigma=sigma, seed=seed)
    return [boxes_on_ground(int(sizes[w]), seed=seed * 100003 + w, sphere_fraction=0.5)
            for w in range(len(sizes))]


def serving_requests(n_requests: int, rate_hz: float, seed: int = 5,
                     lo: int = 8, hi: int = 4096, steps: int = 60) -> List[Dict[str, Any]]:
    """Poisson arrivals (mean ``rate_hz`` per second) of box/sphere worlds with
    log-uniform body counts in [lo, hi], each asking for ``steps`` steps."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, size=n_requests))
    n = np.exp(rng.uniform(np.log(lo), np.log(hi), size=n_requests)).astype(np.int64)
    return [dict(arrival=float(t[i]), steps=steps, n=int(n[i]), seed=seed * 7919 + i)
            for i in range(n_requests)]


def hinge_chains(n_chains: int = 20, links: int = 50, seed: int = 4, spacing: float = 3.0,
                 top: float = 60.0) -> Dict[str, Any]:
    """Chains of 0.2 x 0.2 x 1.0 box links hanging from world anchors, pivots at
    the link ends (each chain starts tilted so it swings).  Joints are listed in
    ``spec['joints']`` as (a, b, anchor) with -1 for the world."""
    rng = np.random.default_rng(seed)
    spec = new_spec()
    _ground(spec, half=(40.0, 0.5, 40.0))
    spec['shapes'].append(('box', (0.2, 1.0, 0.2)))
    link = len(spec['shapes']) - 1
    spec['joints'] = []
    x0 = -0.5 * spacing * (n_chains - 1)
    for c in range(n_chains):
        ax = np.array([x0 + c * spacing, top, 0.0])
        tilt = rng.uniform(-0.6, 0.6)
        d = np.array([np.sin(tilt), -np.cos(tilt), 0.0])
        q = (0.0, 0.0, float(np.sin(0.5 * tilt)), float(np.cos(0.5 * tilt)))
        prev = -1
        for k in range(links):
            centre = ax + d * (k + 0.5)
            spec['bodies'].append(dict(type='dynamic', shape=link, mass=1.0, position=tuple(centre),
                                       orientation=q, material=0))
            idx = len(spec['bodies']) - 1
            spec['joints'].append((prev, idx, tuple(ax + d * k)))
            prev = idx
    return spec


def convex_and_chains(n_hulls: int = 2000, n_chains: int = 20, links: int = 50, seed: int = 4) -> Dict[str, Any]:
    """The convex-and-joints workload: :func:`convex_heap`'s hulls and
    :func:`hinge_chains`' chains (anchored to the world) in one world over one
    ground box of half extents (40, 0.5, 40)."""
    spec = convex_heap(n_hulls, seed=seed)
    spec['shapes'][0] = ('box', (80.0, 1.0, 80.0))          # the heap's ground, enlarged
    chains = hinge_chains(n_chains, links, seed=seed)
    s0, b0 = len(spec['shapes']), len(spec['bodies'])
    # the chains' own ground is body 0 / shape 0: dropped, the rest renumbered
    spec['shapes'] += chains['shapes'][1:]
    for b in chains['bodies'][1:]:
        b = dict(b)
        b['shape'] = b['shape'] - 1 + s0
        spec['bodies'].append(b)
    spec['joints'] = [(a - 1 + b0 if a >= 0 else -1, c - 1 + b0, anchor) for a, c, anchor in chains['joints']]
    return spec


def build(spec: Dict[str, Any], world_cls: Any, model: Any, **extra) -> Any:
    """Build ``spec`` into ``world_cls(**spec['world'], **extra)`` using ``model``'s types."""
    kwargs = dict(spec['world'])
    kwargs.update(extra)
    kwargs.setdefault('gravity', model.Gravity(gravity=9.81, direction=(0.0, -1.0, 0.0)))
    w = world_cls(**kwargs)
    mats = [w.add_material(model.Material(**m)) for m in spec['materials']]
    shapes = []
    for s in spec['shapes']:
        if s[0] == 'box':
            shapes.append(w.add_shape(model.Shape.box(s[1])))
        elif s[0] == 'sphere':
            shapes.append(w.add_shape(model.Shape.sphere(s[1])))
        elif s[0] == 'convex':
            shapes.append(w.add_shape(model.Shape.convex(np.asarray(s[1]))))
        else:
            raise ValueError(s[0])
    for b in spec['bodies']:
        # a body with 'trigger' carries its shape as a sensor (no collider)
        sensor = bool(b.get('trigger', False))
        w.add_body(model.Motion(type=b['type'], mass=b.get('mass', 1.0),
                                linearVelocity=tuple(b.get('linear_velocity', (0.0, 0.0, 0.0))),
                                angularVelocity=tuple(b.get('angular_velocity', (0.0, 0.0, 0.0)))),
                   collider=None if sensor else model.Collider(
                       shape=shapes[b['shape']],
                       physicsMaterial=mats[b['material']] if b.get('material', -1) >= 0 else -1),
                   trigger=model.Trigger(shape=shapes[b['shape']]) if sensor else None,
                   position=tuple(float(x) for x in b['position']),
                   orientation=tuple(float(x) for x in b['orientation']))
    if spec.get('joints'):
        joint_cls = extra_joint_cls or _joint_class(world_cls)
        for a, b, anchor in spec['joints']:
            w.add_joint_constraint(joint_cls(a, b, anchor))
    return w


extra_joint_cls = None


def _joint_class(world_cls: Any) -> Any:
    """The PointConstraint class that belongs with ``world_cls``'s package."""
    import importlib
    pkg = world_cls.__module__.rsplit('.', 1)[0]
    return importlib.import_module(pkg + '.joints').PointConstraint


def arrays(spec: Dict[str, Any], gravity=(0.0, -9.81, 0.0), sleep_enabled: bool = True,
           velocity_iterations: int = 10) -> Dict[str, Any]:
    """``export_world_arrays`` of ``build(spec)`` computed directly with numpy.

    Only specs of box/sphere bodies with one material and the default filter
    (everything the generators above produce); the result is identical to
    building a ``PhysicsWorld`` and exporting it (checked by the tests), but
    fast enough for hundreds of thousands of bodies.
    """
    from .world import KIND_BOX, KIND_SPHERE
    from . import model
    bodies = spec['bodies']
    n = len(bodies)
    shapes = spec['shapes']
    sh = np.array([b['shape'] for b in bodies], dtype=np.int64)
    is_box = np.array([shapes[s][0] == 'box' for s in range(len(shapes))])[sh] if n else np.zeros(0, bool)
    if any(s[0] not in ('box', 'sphere') for s in shapes):
        raise ValueError('arrays() handles box and sphere specs only; use build()')
    size = np.zeros((len(shapes), 3))
    rad = np.zeros(len(shapes))
    for i, s in enumerate(shapes):
        if s[0] == 'box':
            size[i] = s[1]
        else:
            rad[i] = s[1]
    dyn = np.array([b['type'] == 'dynamic' for b in bodies])
    mt = np.array([{'static': 0, 'kinematic': 1, 'dynamic': 2}[b['type']] for b in bodies], np.int32)
    mass = np.array([max(b.get('mass', 1.0), 1e-9) for b in bodies])
    half = np.where(is_box[:, None], 0.5 * size[sh], rad[sh][:, None])
    x, y, z = size[sh, 0], size[sh, 1], size[sh, 2]
    k = mass / 12.0
    d_box = np.stack([k * (y * y + z * z), k * (x * x + z * z), k * (x * x + y * y)], axis=1)
    v = 0.4 * mass * rad[sh] ** 2
    d_sph = np.stack([v, v, v], axis=1)
    diag = np.where(is_box[:, None], d_box, d_sph)
    with np.errstate(divide='ignore'):
        inv_inertia = np.where(diag > 0, 1.0 / np.maximum(diag, 1e-12), 0.0)
    inv_inertia[~dyn] = 0.0
    inv_mass = np.where(dyn, 1.0 / mass, 0.0)
    pos = np.array([b['position'] for b in bodies], dtype='d').reshape(n, 3)
    quat = np.array([b['orientation'] for b in bodies], dtype='d').reshape(n, 4)
    lv = np.array([b.get('linear_velocity', (0.0, 0.0, 0.0)) for b in bodies], dtype='d').reshape(n, 3)
    av = np.array([b.get('angular_velocity', (0.0, 0.0, 0.0)) for b in bodies], dtype='d').reshape(n, 3)
    mats = spec['materials']
    if len(mats) != 1 or any(b.get('material', 0) != 0 for b in bodies):
        raise ValueError('arrays() handles single-material specs only; use build()')
    m0 = model.Material(**mats[0])
    fr = model.combine(m0.dynamicFriction, m0.dynamicFriction, m0.frictionCombine)
    rs = model.combine(m0.restitution, m0.restitution, m0.restitutionCombine)
    zeros = np.zeros(n)
    g = np.asarray(gravity, dtype='d')
    return dict(
        n=n, position=pos, orientation=quat, prev_position=pos.copy(), prev_orientation=quat.copy(),
        linear_velocity=lv, angular_velocity=av, inv_mass=inv_mass, inv_inertia=inv_inertia,
        gravity_factor=np.ones(n), linear_damping=zeros, angular_damping=zeros.copy(),
        quadratic_drag=zeros.copy(), sleep_timer=zeros.copy(), motion_type=mt,
        collider_shape=sh.astype(np.int32), trigger_shape=np.full(n, -1, np.int32),
        awake=np.ones(n, np.uint8), kind=np.where(is_box, KIND_BOX, KIND_SPHERE).astype(np.int32),
        half=np.ascontiguousarray(half), hull_off=np.full(n, -1, np.int32), hull_cnt=np.zeros(n, np.int32),
        hull_pts=np.zeros((0, 3)), mat_slot=np.zeros(n, np.int32), n_mat=1,
        friction_table=np.array([[fr]]), restitution_table=np.array([[rs]]),
        filt_slot=np.zeros(n, np.int32), n_filt=1, filter_table=np.ones((1, 1), np.uint8),
        gravity=g, default_linear_damping=0.0, default_angular_damping=0.0,
        sleep_enabled=int(sleep_enabled), velocity_iterations=int(velocity_iterations), has_collision=1,
        n_joints=0, joint_iterations=10, joint_a=np.zeros(0, np.int32), joint_b=np.zeros(0, np.int32),
        joint_anchor=np.zeros((0, 3)), joint_local_a=np.zeros((0, 3)), joint_local_b=np.zeros((0, 3)),
        joint_fixed=np.zeros(0, np.uint8))
