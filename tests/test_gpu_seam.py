"""The backend seam on the GPU.

Stage path: a host-column world that calls the backend in the reference's
``PhysicsWorld.step`` order (``tests/seam_world.py``) runs on ``HipBackend``
and equals the oracle bit for bit, with trigger events dispatched inside the
collision stage and the warm-start cache kept across a world edit.  Stage
entry points (integrate forces / positions, sleep) equal the oracle's
restatements.  Resident path: host edits between steps (parameters, state, an
added body, a joint world reload) reach the device and leave the world equal
to the oracle given the same edits.
"""
import sys

import numpy as np
import pytest

import oracle
from omi_physics_amd import PhysicsWorld, model, scenes, _ffi
from omi_physics_amd._ffi import _d, _i32, _u8, check, ptr, c64, ci32
from omi_physics_amd.backend import HipBackend, HipCollisionStages
from tests import helpers
from tests.seam_world import SeamWorld

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0
KEYS = ('position', 'orientation', 'linear_velocity', 'angular_velocity', 'sleep_timer')


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_state(cols, o, where):
    for key in KEYS:
        if not np.array_equal(cols[key], o[key]):
            raise AssertionError('%s: %s differs (max %.3g)' % (where, key, np.abs(cols[key] - o[key]).max()))
    assert np.array_equal(np.asarray(cols['awake'], np.uint8), o['awake']), where


def seam_cols(w):
    return {k: getattr(w, k) for k in KEYS + ('awake',)}


def test_reference_order_world_runs_on_the_stage_path():
    """A pile falling through sensors, stepped by a host world in the reference's
    call order on HipBackend: the host collision stub is replaced by the device
    stage, and 120 FP64 steps equal the oracle bit for bit (states, contacts,
    wake and sleep).  Trigger events arrive during the collision stage: at each
    event the world's positions are the pre-integration ones, x_end = x_event + v dt."""
    spec = helpers.trigger_scene(50, seed=6)
    w = scenes.build(spec, SeamWorld, model)
    w.backend = HipBackend()
    arrays = scenes.build(spec, PhysicsWorld, model).export_arrays()
    ow = oracle.OracleWorld(arrays)
    seen = []

    def listener(kind, t, o):
        seen.append((kind, t, o, w.position.copy(), w.linear_velocity.copy()))
    w.add_trigger_listener(listener)
    events = 0
    for s in range(120):
        seen.clear()
        w.step(DT)
        ow.step(DT)
        assert isinstance(w._collision, HipCollisionStages)
        same_state(seam_cols(w), ow.state(), 'stage path step %d' % s)
        oc = ow.contacts()
        assert len(w.contacts) == len(oc['a'])
        for k, c in enumerate(w.contacts[:50]):
            assert (c.a, c.b) == (oc['a'][k], oc['b'][k]) and c.normal_impulse == oc['normal_impulse'][k]
        for kind, t, o, x_event, v_event in seen:
            moving = (w.motion_type == 2) & w.awake
            np.testing.assert_array_equal(x_event[moving] + v_event[moving] * DT, w.position[moving])
        events += len(seen)
    say('stage path: 120 steps bit-exact, %d trigger events' % events)
    assert events > 20


def test_stage_path_keeps_the_warm_start_cache_across_an_added_body():
    """Adding a far static body after 40 steps reloads the stage's device world with
    its warm-start cache: the next 60 steps equal an oracle world that had the body
    from the start (a lost cache would change the impulses)."""
    spec = helpers.pile(90, seed=8)
    far = dict(type='static', shape=1, position=(200.0, 0.5, 200.0), orientation=(0.0, 0.0, 0.0, 1.0), material=0)
    full = dict(spec, bodies=spec['bodies'] + [far])
    w = scenes.build(spec, SeamWorld, model)
    w.backend = HipBackend()
    ow = oracle.OracleWorld(scenes.build(full, PhysicsWorld, model).export_arrays())
    for s in range(100):
        if s == 40:
            w.add_body(model.Motion(type=model.STATIC), collider=model.Collider(shape=1, physicsMaterial=0),
                       position=far['position'])
        w.step(DT)
        ow.step(DT)
        if s >= 40:
            same_state(seam_cols(w), ow.state(), 'after the edit, step %d' % s)
    assert len(w.contacts) > 50


def test_stage_entry_points_equal_the_oracle():
    """omi_integrate_forces (per-body gravity, damping, drag), omi_integrate_positions
    and omi_sleep on random bodies and contacts, bit for bit against the oracle."""
    L = _ffi.load()
    rng = np.random.default_rng(4)
    n = 3000
    mt = rng.choice([0, 1, 2], size=n, p=[0.1, 0.1, 0.8]).astype(np.int32)
    aw = (rng.uniform(size=n) < 0.8).astype(np.uint8)
    g = rng.normal(size=(n, 3)) * 5
    gf, ld, ad, qd = rng.uniform(0, 2, n), rng.uniform(0, 3, n), rng.uniform(0, 3, n), rng.uniform(0, 0.5, n)
    lv, av = rng.normal(size=(n, 3)) * 3, rng.normal(size=(n, 3)) * 2
    glv, gav = lv.copy(), av.copy()
    check(L.omi_integrate_forces(n, ptr(mt, _i32), ptr(aw, _u8), ptr(c64(g), _d), *[ptr(c64(x), _d) for x in (gf, ld, ad, qd)],
                                 0.07, 0.11, DT, ptr(glv, _d), ptr(gav, _d)))
    olv, oav = lv.copy(), av.copy()
    oracle.integrate_forces(mt, aw, g, gf, ld, ad, qd, 0.07, 0.11, olv, oav, DT)
    assert np.array_equal(glv, olv) and np.array_equal(gav, oav)
    assert not np.array_equal(glv, lv)
    pos = rng.normal(size=(n, 3)) * 10
    q = rng.normal(size=(n, 4)); q /= np.linalg.norm(q, axis=1, keepdims=True)
    gp, gq, op_, oq = pos.copy(), q.copy(), pos.copy(), q.copy()
    check(L.omi_integrate_positions(n, ptr(mt, _i32), ptr(aw, _u8), ptr(c64(glv), _d), ptr(c64(gav), _d), DT,
                                    ptr(gp, _d), ptr(gq, _d)))
    oracle.integrate_positions(mt, aw, glv, gav, op_, oq, DT)
    assert np.array_equal(gp, op_) and np.array_equal(gq, oq)
    # sleep: slow bodies with grown timers, contacts forming islands of various readiness
    lv = rng.normal(size=(n, 3)) * np.where(rng.uniform(size=(n, 1)) < 0.6, 0.02, 1.0)
    av = rng.normal(size=(n, 3)) * np.where(rng.uniform(size=(n, 1)) < 0.6, 0.05, 1.0)
    tm = rng.uniform(0, 0.6, n)
    K = 4000
    ca, cb = rng.integers(0, n, K).astype(np.int32), rng.integers(0, n, K).astype(np.int32)
    keep = ca != cb
    ca, cb = ca[keep], cb[keep]
    g_ = [lv.copy(), av.copy(), aw.copy(), tm.copy()]
    check(L.omi_sleep(n, len(ca), ptr(ca, _i32), ptr(cb, _i32), ptr(mt, _i32), DT, 0.08, 0.25, 0.4, ptr(g_[0], _d),
                      ptr(g_[1], _d), ptr(g_[2], _u8), ptr(g_[3], _d)))
    o_ = [lv.copy(), av.copy(), aw.copy(), tm.copy()]
    oracle.update_sleep(ca, cb, mt, *o_, DT)
    for x, y in zip(g_, o_):
        assert np.array_equal(x, y)
    assert ((aw == 1) & (g_[2] == 0)).any() and ((aw == 0) & (g_[2] == 1)).any()   # bodies slept and woke


def test_resident_world_keeps_its_cache_across_an_added_body():
    """The resident PhysicsWorld reloads its device world when a body is added after
    stepping and carries the warm-start cache over: bit-equal to an oracle world that
    had the far body from the start."""
    spec = helpers.pile(100, seed=9)
    far = dict(type='static', shape=1, position=(-300.0, 0.5, 5.0), orientation=(0.0, 0.0, 0.0, 1.0), material=0)
    w = scenes.build(spec, PhysicsWorld, model)
    ow = oracle.OracleWorld(scenes.build(dict(spec, bodies=spec['bodies'] + [far]), PhysicsWorld, model).export_arrays())
    for s in range(90):
        if s == 30:
            w.add_body(model.Motion(type=model.STATIC), collider=model.Collider(shape=1, physicsMaterial=0),
                       position=far['position'])
        w.step(DT)
        ow.step(DT)
        if s >= 30 and s % 10 == 9:
            st = ow.state()
            for key in KEYS:
                assert np.array_equal(getattr(w, key), st[key]), (s, key)
    assert len(w.contacts) > 50


def test_resident_world_sees_parameter_and_state_edits():
    """Edits between steps — gravity factors, damping, drag, inverse mass, world
    gravity and default damping, and a body's teleport and velocity — reach the
    device: the world stays bit-equal to the oracle given the same edits."""
    arrays_spec = helpers.pile(80, seed=10)
    w = scenes.build(arrays_spec, PhysicsWorld, model)
    ow = oracle.OracleWorld(w.export_arrays())
    rng = np.random.default_rng(2)
    n = w.body_count
    for s in range(120):
        if s == 20:
            gf = w.gravity_factor; gf[1:] = rng.uniform(0.5, 1.5, n - 1)
            w.linear_damping[1:] = rng.uniform(0, 0.5, n - 1)
            w.quadratic_drag[5:20] = 0.2
            w.angular_damping[3:9] = 0.4
            ow.set_params(gravity_factor=w._raw('gravity_factor'), linear_damping=w._raw('linear_damping'),
                          quadratic_drag=w._raw('quadratic_drag'), angular_damping=w._raw('angular_damping'))
        if s == 50:
            w.gravity = model.Gravity(gravity=4.0, direction=(0.3, -1.0, 0.0))
            w.default_linear_damping, w.default_angular_damping = 0.1, 0.2
            w.inv_mass[10] = 0.25
            ow.set_params(gravity=w.resolve_gravity(), default_damping=(0.1, 0.2), inv_mass=w._raw('inv_mass'))
        if s == 80:
            pos = w.position; pos[7] = (0.0, 9.0, 0.0)
            w.linear_velocity[7] = (0.0, -3.0, 0.0)
            w.wake(7)
            ow.set_state(position=w._raw('position'), linear_velocity=w._raw('linear_velocity'),
                         awake=w._raw('awake'), sleep_timer=w._raw('sleep_timer'))
        w.step(DT)
        ow.step(DT)
        if s % 10 == 9:
            st = ow.state()
            for key in KEYS:
                assert np.array_equal(getattr(w, key), st[key]), (s, key)


def test_joint_anchors_given_and_kept_across_a_reload():
    """PointConstraint with an explicit body-frame anchor equals the oracle given the
    same anchor; a chain world reloaded after 30 steps (a far body added) keeps the
    anchors the device fixed at the first step, bit-equal to the oracle world that had
    the body from the start."""
    from omi_physics_amd.joints import PointConstraint
    w = PhysicsWorld(gravity=model.Gravity(gravity=9.81), sleep_enabled=False)
    s = w.add_shape(model.Shape.box((0.2, 1.0, 0.2)))
    bob = w.add_body(model.Motion(mass=1.0), collider=model.Collider(shape=s), position=(0.8, 2.0, 0.0))
    w.add_joint_constraint(PointConstraint(-1, bob, (0.0, 2.5, 0.0), local_b=(0.0, 0.5, 0.0)))
    ow = oracle.OracleWorld(w.export_arrays())
    for _ in range(90):
        w.step(DT)
    ow.step(DT, 90)
    assert np.array_equal(w.position, ow.state()['position'])
    spec = scenes.hinge_chains(n_chains=3, links=8, seed=5, spacing=1.5, top=10.0)
    far = dict(type='static', shape=1, position=(300.0, 0.5, 0.0), orientation=(0.0, 0.0, 0.0, 1.0), material=0)
    w = scenes.build(spec, PhysicsWorld, model)
    ow = oracle.OracleWorld(scenes.build(dict(spec, bodies=spec['bodies'] + [far]), PhysicsWorld,
                                         model).export_arrays())
    for s in range(70):
        if s == 30:
            w.add_body(model.Motion(type=model.STATIC), collider=model.Collider(shape=1, physicsMaterial=0),
                       position=far['position'])
        w.step(DT)
        ow.step(DT)
    st = ow.state()
    for key in KEYS:
        assert np.array_equal(getattr(w, key), st[key]), key
