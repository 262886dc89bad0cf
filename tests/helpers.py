"""Shared scene builders for the tests (the same specs the golden script used)."""
import numpy as np

from omi_physics_amd import PhysicsWorld, model, scenes, mathutil


def traj_spec(name):
    """The small scenes of tests/golden/traj_*.npz (see make_golden.py)."""
    if name == 'dropbox':
        spec = scenes.new_spec(); scenes._ground(spec); spec['shapes'].append(('box', (1.0, 1.0, 1.0)))
        spec['bodies'].append(dict(type='dynamic', shape=1, mass=1.0, position=(0.3, 3.0, -0.2),
                                   orientation=tuple(mathutil.quat_from_axis_angle((1, 2, 3), 0.7)), material=0))
    elif name == 'stack5':
        spec = scenes.new_spec(); scenes._ground(spec); spec['shapes'].append(('box', (1.0, 1.0, 1.0)))
        for k in range(5):
            spec['bodies'].append(dict(type='dynamic', shape=1, mass=1.0, position=(0.0, 0.5 + k * 1.001, 0.0),
                                       orientation=(0.0, 0.0, 0.0, 1.0), material=0))
    elif name == 'rolling':
        spec = scenes.new_spec(friction=0.3); scenes._ground(spec); spec['shapes'].append(('sphere', 0.5))
        spec['bodies'].append(dict(type='dynamic', shape=1, mass=2.0, position=(-3.0, 0.6, 0.0),
                                   orientation=(0.0, 0.0, 0.0, 1.0), material=0, linear_velocity=(2.0, 0.0, 0.5)))
    elif name == 'mixed24':
        spec = scenes.boxes_on_ground(24, seed=11, spread=2.0, y_range=(1.0, 8.0), sphere_fraction=0.5)
    else:
        raise KeyError(name)
    return spec


def world(spec, **kw):
    return scenes.build(spec, PhysicsWorld, model, **kw)


def sap_filter_table(filt):
    """Slots and table for the filters used by tests/golden/sap.npz scenes 2-3:
    0 = systems ('a',) excluding ('b',); 1 = ('b',); 2 = ('c',) colliding only
    with ('c',); -1 = the default filter."""
    filters = [model.CollisionFilter(collisionSystems=('a',), notCollideWithSystems=('b',)),
               model.CollisionFilter(collisionSystems=('b',)),
               model.CollisionFilter(collisionSystems=('c',), collideWithSystems=('c',))]
    def get(i):
        return filters[i] if 0 <= i < len(filters) else model.DEFAULT_FILTER
    vals, slots = np.unique(filt, return_inverse=True)
    tab = np.array([[get(int(a)).collides_with(get(int(b))) for b in vals] for a in vals], np.uint8)
    return slots.astype(np.int32), tab


def pile(n=120, seed=3, spheres=0.5, spread=2.0, y=(0.6, 10.0)):
    """A dense pile that produces many contacts within a few dozen steps."""
    return scenes.boxes_on_ground(n, seed=seed, spread=spread, y_range=y, sphere_fraction=spheres)


def trigger_scene(n=60, seed=4):
    """A pile falling through sensors: static box and sphere trigger zones on and above
    the ground, a rotated box zone, and a dynamic sphere that is itself a trigger."""
    spec = scenes.boxes_on_ground(n, seed=seed, spread=3.0, y_range=(1.0, 9.0), sphere_fraction=0.5)
    spec['shapes'] += [('box', (3.0, 2.0, 3.0)), ('sphere', 1.5), ('box', (6.0, 0.6, 1.2)), ('sphere', 0.4)]
    zb, zs, zr, ts = range(len(spec['shapes']) - 4, len(spec['shapes']))
    ident = (0.0, 0.0, 0.0, 1.0)
    rot = tuple(mathutil.quat_from_axis_angle((0.3, 1.0, 0.2), 0.7))
    spec['bodies'] += [
        dict(type='static', shape=zb, position=(-1.0, 1.5, 0.5), orientation=ident, trigger=True),
        dict(type='static', shape=zs, position=(1.5, 4.0, -1.0), orientation=ident, trigger=True),
        dict(type='static', shape=zr, position=(0.0, 2.5, 0.0), orientation=rot, trigger=True),
        dict(type='dynamic', shape=ts, position=(0.2, 6.0, 0.1), orientation=ident, trigger=True,
             linear_velocity=(0.0, -1.0, 0.0)),
    ]
    return spec


def damped_kinematic_scene(n=90, seed=21, asleep_every=4, spread=3.0):
    """A pile whose bodies carry every integrate_forces term and which kinematic
    pushers sweep through, with part of it starting asleep.

    Per-body linear/angular damping, quadratic drag and gravity factors are
    random; the world adds default damping; three kinematic boxes move at
    constant velocity (one spinning) across the pile; every ``asleep_every``-th
    dynamic body starts asleep, so contacts with an awake mover wake it
    (solver.py:321-337) and the kinematic bodies count as awake movers in the
    solid-pair filter (world.py:381-384).  The dynamic bodies start within
    ``spread`` of the centre in x and z.  Returns ``export_world_arrays`` output.
    """
    from omi_physics_amd import PhysicsWorld
    rng = np.random.default_rng(seed)
    w = PhysicsWorld(gravity=model.Gravity(gravity=9.81, direction=(0.0, -1.0, 0.0)),
                     default_linear_damping=0.05, default_angular_damping=0.12)
    mat = w.add_material(model.Material(staticFriction=0.5, dynamicFriction=0.5))
    ground = w.add_shape(model.Shape.box((24.0, 1.0, 24.0)))
    w.add_body(model.Motion(type=model.STATIC), collider=model.Collider(shape=ground, physicsMaterial=mat),
               position=(0.0, -0.5, 0.0))
    box = w.add_shape(model.Shape.box((1.0, 1.0, 1.0)))
    ball = w.add_shape(model.Shape.sphere(0.5))
    pusher = w.add_shape(model.Shape.box((1.0, 1.6, 5.0)))
    for k in range(n):
        q = mathutil.quat_from_axis_angle(tuple(rng.normal(size=3)), float(rng.uniform(0, 2 * np.pi)))
        m = model.Motion(mass=float(rng.uniform(0.5, 2.0)),
                         linearVelocity=tuple(rng.normal(scale=0.5, size=3)),
                         angularVelocity=tuple(rng.normal(scale=1.0, size=3)),
                         gravityFactor=float(rng.uniform(0.5, 1.5)),
                         linearDamping=float(rng.uniform(0.0, 0.8)),
                         angularDamping=float(rng.uniform(0.0, 0.8)),
                         quadraticDrag=float(rng.uniform(0.0, 0.3)))
        w.add_body(m, collider=model.Collider(shape=box if k % 2 else ball, physicsMaterial=mat),
                   position=(float(rng.uniform(-spread, spread)), float(rng.uniform(0.6, 5.0)),
                             float(rng.uniform(-spread, spread))),
                   orientation=tuple(q))
    for x0, v, spin in ((-8.0, (2.0, 0.0, 0.0), (0.0, 0.0, 0.0)), (8.0, (-1.5, 0.0, 0.3), (0.0, 0.8, 0.0)),
                        (0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))):
        w.add_body(model.Motion(type=model.KINEMATIC, linearVelocity=v, angularVelocity=spin),
                   collider=model.Collider(shape=pusher, physicsMaterial=mat),
                   position=(x0, 0.8, 0.0) if x0 else (0.0, 0.8, 9.0))
    arrays = w.export_arrays()
    aw = arrays['awake'].copy()
    dyn = np.flatnonzero(arrays['motion_type'] == 2)
    aw[dyn[::asleep_every]] = 0
    arrays['awake'] = aw
    return arrays
