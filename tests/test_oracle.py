"""The oracle pinned against the reference's golden vectors, closed forms and brute force (CPU)."""
import numpy as np
import pytest

import oracle
from tests import helpers


def test_refit_bit_exact_with_reference(golden):
    g = golden('refit.npz')
    lo, hi = oracle.refit(g['pos'], g['quat'], g['kind'], g['half'])
    assert np.array_equal(lo, g['lo']) and np.array_equal(hi, g['hi'])


def test_refit_box_closed_form():
    # axis-aligned box: AABB = centre ± half ± margin; 90° about z swaps x/y extents
    q = np.array([[0, 0, 0, 1.0], [0, 0, np.sin(np.pi / 4), np.cos(np.pi / 4)]])
    pos = np.array([[1.0, 2.0, 3.0], [0.0, 0.0, 0.0]])
    lo, hi = oracle.refit(pos, q, [2, 2], [[0.5, 1.0, 1.5], [0.5, 1.0, 1.5]], margin=0.05)
    assert np.allclose(lo[0], [0.45, 0.95, 1.45]) and np.allclose(hi[0], [1.55, 3.05, 4.55])
    assert np.allclose(hi[1], [1.05, 0.55, 1.55], atol=1e-12)


@pytest.mark.parametrize('scene', range(4))
def test_sap_order_matches_reference(golden, scene):
    g = golden('sap.npz')
    filt = g['filt%d' % scene]
    slots, tab = helpers.sap_filter_table(filt)
    pairs, _ = oracle.sap_pairs(g['lo%d' % scene], g['hi%d' % scene], g['mt%d' % scene], slots, tab)
    assert np.array_equal(pairs, g['pairs%d' % scene])


def test_sap_equals_brute_force():
    rng = np.random.default_rng(1)
    n = 200
    c = rng.uniform(-3, 3, size=(n, 3)); h = rng.uniform(0.1, 0.6, size=(n, 3))
    lo, hi = c - h, c + h
    mt = rng.choice([0, 1, 2], size=n, p=[0.2, 0.1, 0.7]).astype(np.int32)
    pairs, _ = oracle.sap_pairs(lo, hi, mt)
    brute = {(i, j) for i in range(n) for j in range(i + 1, n)
             if (mt[i] == 2 or mt[j] == 2) and np.all(lo[i] <= hi[j]) and np.all(lo[j] <= hi[i])}
    assert set(map(tuple, pairs.tolist())) == brute
    assert len(pairs) == len(brute)


def test_box_box_matches_reference(golden):
    g = golden('boxbox.npz')
    for i in range(len(g['ca'])):
        n, nrm, pts, dep = oracle.box_box(g['ca'][i], g['Ra'][i], g['ha'][i], g['cb'][i], g['Rb'][i], g['hb'][i])
        assert n == g['count'][i]
        if n:
            np.testing.assert_allclose(nrm, g['normal'][i], atol=1e-12)
            np.testing.assert_allclose(pts, g['points'][i][:n], atol=1e-12)
            np.testing.assert_allclose(dep, g['depths'][i][:n], atol=1e-12)


def test_box_box_closed_forms():
    I = np.eye(3)
    # reference test_collide_primitives.py:53-60: overlap 0.5 on x
    n, nrm, pts, dep = oracle.box_box([0, 0, 0], I, [1, 1, 1], [1.5, 0, 0], I, [1, 1, 1])
    assert n >= 1 and np.allclose(np.abs(nrm), [1, 0, 0]) and abs(dep.max() - 0.5) < 1e-12
    # resting crate: four face contacts, depth 0.01, normal +y
    n, nrm, pts, dep = oracle.box_box([0, 0, 0], I, [5, 0.5, 5], [0, 0.99, 0], I, [0.5, 0.5, 0.5])
    assert n == 4 and np.allclose(nrm, [0, 1, 0]) and np.allclose(dep, 0.01)
    assert np.allclose(pts[:, 1], 0.5)        # seated on the reference face
    n, *_ = oracle.box_box([0, 0, 0], I, [1, 1, 1], [3, 0, 0], I, [1, 1, 1])
    assert n == 0


def test_spheres_match_reference(golden):
    g = golden('spheres.npz')
    for i in range(len(g['sa'])):
        n, nrm, pt, dep = oracle.sphere_sphere(g['sa'][i], g['ra'][i], g['sb'][i], g['rb'][i])
        e = g['ss'][i]
        assert n == int(e[0])
        if n:
            np.testing.assert_allclose(np.r_[nrm, pt[0], dep], e[1:], atol=1e-12)
        n, nrm, pt, dep = oracle.sphere_box(g['sc'][i], g['sr'][i], g['bc'][i], g['bR'][i], g['bh'][i])
        e = g['sbx'][i]
        assert n == int(e[0])
        if n:
            np.testing.assert_allclose(np.r_[nrm, pt[0], dep], e[1:], atol=1e-12)


def test_sphere_closed_forms():
    n, nrm, pt, dep = oracle.sphere_sphere([0, 0, 0], 1.0, [1.5, 0, 0], 1.0)
    assert n == 1 and np.allclose(nrm, [1, 0, 0]) and abs(dep[0] - 0.5) < 1e-15 and abs(pt[0][0] - 0.75) < 1e-15
    n, nrm, pt, dep = oracle.sphere_box([0, 1.3, 0], 0.5, [0, 0, 0], np.eye(3), [1, 1, 1])
    assert n == 1 and np.allclose(nrm, [0, -1, 0]) and abs(dep[0] - 0.2) < 1e-12


def _hull(g, i):
    return g['hull_pts'][g['hull_off'][i]:g['hull_off'][i + 1]]


def test_gjk_epa_matches_reference(golden):
    g = golden('gjk.npz')
    for k in range(len(g['idx'])):
        i, j = g['idx'][k]
        kb, hb = (2, np.array([0.8, 0.3, 0.6])) if k % 10 == 9 else (3, None)
        args = (3, g['pos'][k, 0], g['R'][k, 0], None, _hull(g, i), kb, g['pos'][k, 1], g['R'][k, 1], hb, _hull(g, j))
        assert oracle.gjk_intersect(*args) == bool(g['res'][k, 0])
        n, nrm, pt, dep = oracle.collide_convex(*args)
        e = g['res'][k]
        assert n == int(e[1])
        if n:
            np.testing.assert_allclose(nrm, e[2:5], atol=1e-9)
            np.testing.assert_allclose(pt, e[5:8], atol=1e-9)
            assert abs(dep - e[8]) < 1e-9


def test_epa_closed_form():
    # reference test_gjk_epa.py:31-38: unit cubes overlapping 0.3 on x
    cube = np.array([(sx, sy, sz) for sx in (-.5, .5) for sy in (-.5, .5) for sz in (-.5, .5)])
    n, nrm, pt, dep = oracle.collide_convex(3, [0, 0, 0], np.eye(3), None, cube, 3, [0.7, 0, 0], np.eye(3), None, cube)
    assert n == 1 and abs(dep - 0.3) < 1e-9 and np.allclose(np.abs(nrm), [1, 0, 0], atol=1e-9)


def test_solver_matches_reference(golden):
    g = golden('solver.npz')
    for isl in range(12):
        pos = g['pos_in%d' % isl].copy(); lv = g['lv_in%d' % isl].copy(); av = g['av_in%d' % isl].copy()
        ni = g['seedn%d' % isl].copy(); ti = g['seedt%d' % isl].copy()
        oracle.prepare_and_solve(g['a%d' % isl], g['b%d' % isl], g['point%d' % isl], g['normal%d' % isl],
                                 g['depth%d' % isl], pos, lv, av, g['inv_mass_in%d' % isl], g['invIw%d' % isl],
                                 g['rest%d' % isl], g['fric%d' % isl], ni, ti)
        for got, key in ((pos, 'pos_out'), (lv, 'lv_out'), (av, 'av_out'), (ni, 'nimp_out'), (ti, 'timp_out')):
            np.testing.assert_allclose(got, g['%s%d' % (key, isl)], atol=1e-11, rtol=0)


def test_solver_single_contact_closed_form():
    # a unit-mass sphere resting on static ground: one normal impulse = m g dt
    g_dt = 9.81 / 60.0
    pos = np.array([[0, 0, 0], [0, 0.5, 0]], float)
    lv = np.array([[0, 0, 0], [0, -g_dt, 0]], float); av = np.zeros((2, 3))
    invIw = np.zeros((2, 3, 3)); invIw[1] = np.eye(3) * 10.0
    ni = np.zeros(1); ti = np.zeros((1, 2))
    oracle.prepare_and_solve([0], [1], [[0, 0, 0]], [[0, 1, 0]], [0.0], pos, lv, av, [0.0, 1.0], invIw,
                             [0.0], [0.5], ni, ti, warm_start=False)
    assert abs(ni[0] - g_dt) < 1e-12 and abs(lv[1, 1]) < 1e-12 and np.all(lv[0] == 0)


@pytest.mark.parametrize('name,tol', [('dropbox', 1e-12), ('stack5', 1e-9), ('rolling', 1e-12), ('mixed24', 1e-9)])
def test_trajectory_matches_reference(golden, name, tol):
    g = golden('traj_%s.npz' % name)
    ow = oracle.OracleWorld(helpers.world(helpers.traj_spec(name)).export_arrays())
    s = 0
    for idx, st in enumerate(g['steps']):
        ow.step(1.0 / 60.0, int(st) - s); s = int(st)
        S = ow.state()
        np.testing.assert_allclose(S['position'], g['pos'][idx], atol=tol, rtol=0)
        np.testing.assert_allclose(S['linear_velocity'], g['lv'][idx], atol=100 * tol, rtol=0)
        assert np.array_equal(S['awake'].astype(bool), g['awake'][idx])
        assert ow.counters()['contacts'] == g['ncontacts'][idx]


def test_freefall_closed_form():
    # reference test_physics_integrate.py:16-26: x_n = -g dt² n(n+1)/2
    from omi_physics_amd import PhysicsWorld, model
    w = PhysicsWorld(gravity=model.Gravity(gravity=9.81, direction=(0, -1, 0)))
    w.add_body(model.Motion(type=model.DYNAMIC), position=(0, 0, 0))
    ow = oracle.OracleWorld(w.export_arrays())
    dt, n = 1.0 / 240.0, 240
    ow.step(dt, n)
    assert ow.state()['position'][0, 1] == pytest.approx(-0.5 * 9.81 * dt * dt * n * (n + 1), rel=1e-9)


def test_sleep_settles_a_row_of_boxes():
    # reference test_physics_benchmark.py:61-74
    from omi_physics_amd import PhysicsWorld, model
    w = PhysicsWorld(gravity=model.Gravity(gravity=9.81, direction=(0, -1, 0)))
    box = w.add_shape(model.Shape.box((1, 1, 1)))
    w.add_body(model.Motion(type=model.STATIC), collider=model.Collider(shape=box), position=(0, -0.5, 0))
    for k in range(20):
        w.add_body(model.Motion(type=model.DYNAMIC), collider=model.Collider(shape=box), position=(k * 2.0, 0.5, 0))
    ow = oracle.OracleWorld(w.export_arrays())
    ow.step(1 / 120, 240)
    assert int(ow.state()['awake'][1:].sum()) < 20


def test_mutation_is_caught():
    """A plausible mistake planted in the oracle's inputs (a flipped normal) fails the golden check."""
    g = np.load(__import__('os').path.join(__import__('tests.conftest', fromlist=['GOLDEN']).GOLDEN, 'solver.npz'))
    pos = g['pos_in0'].copy(); lv = g['lv_in0'].copy(); av = g['av_in0'].copy()
    ni = g['seedn0'].copy(); ti = g['seedt0'].copy()
    oracle.prepare_and_solve(g['a0'], g['b0'], g['point0'], -g['normal0'], g['depth0'], pos, lv, av,
                             g['inv_mass_in0'], g['invIw0'], g['rest0'], g['fric0'], ni, ti)
    assert np.abs(lv - g['lv_out0']).max() > 1e-3


def test_solve3_matches_lapack():
    """3x3 LU with partial pivoting against LAPACK, including systems whose first pivot is zero
    or smallest (row swaps in both elimination steps)."""
    rng = np.random.default_rng(5)
    systems = [(np.array([[0.0, 2.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 3.0]]), np.array([1.0, 2.0, 3.0])),
               (np.array([[1e-9, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 5.0, 1.0]]), np.array([1.0, 0.0, -1.0]))]
    for _ in range(2000):
        K = rng.normal(size=(3, 3)) + 3 * np.eye(3) * rng.uniform(0, 1)
        systems.append((K, rng.normal(size=3)))
    for K, r in systems:
        np.testing.assert_allclose(oracle.solve3(K, r), np.linalg.solve(K, r), rtol=1e-9, atol=1e-11)


def test_joint_chain_matches_reference(golden):
    """PointConstraint chains (joints.py:60-121): within 1e-7 m for 60 steps.  Adjacent links start
    end-face to end-face at depth ~0, so a last-bit difference (LAPACK/BLAS vs the restatement) can
    keep or drop a zero-depth clip point: contact counts agree within 2."""
    from omi_physics_amd import scenes
    g = golden('traj_chain.npz')
    ow = oracle.OracleWorld(helpers.world(scenes.hinge_chains(n_chains=3, links=8, seed=5, spacing=1.5,
                                                               top=10.0)).export_arrays())
    s = 0
    for idx, st in enumerate(g['steps']):
        ow.step(1.0 / 60.0, int(st) - s); s = int(st)
        assert abs(ow.counters()['contacts'] - g['ncontacts'][idx]) <= 2, st
        if st <= 60:
            np.testing.assert_allclose(ow.state()['position'], g['pos'][idx], atol=1e-7, rtol=0)


def test_pendulum_keeps_its_length():
    """Invariant: a ball joint to the world holds the anchor distance (Baumgarte-corrected)."""
    from omi_physics_amd import PhysicsWorld, model
    from omi_physics_amd.joints import PointConstraint
    w = PhysicsWorld(gravity=model.Gravity(gravity=9.81, direction=(0, -1, 0)), sleep_enabled=False)
    b = w.add_body(model.Motion(mass=1.0), position=(1.0, 5.0, 0.0))
    w.add_joint_constraint(PointConstraint(-1, b, (0.0, 5.0, 0.0)))
    ow = oracle.OracleWorld(w.export_arrays())
    for _ in range(120):
        ow.step(1.0 / 60.0)
        p = ow.state()['position'][0]
        assert abs(np.linalg.norm(p - [0.0, 5.0, 0.0]) - 1.0) < 0.02


def test_snapshot_axis_angle_closed_forms():
    """ow_snapshot against what the rotation fixes: a quaternion built from a known
    unit axis and angle in [0.1, 2 pi - 0.1] returns that axis and angle (1e-12); any positive
    scaling of the quaternion gives the same result; the identity (either sign) and a
    rotation whose w rounds to 1 give the canonical (0, 1, 0, 0); a small rotation
    keeps its axis; pi about
    z is (0, 0, 1, pi); positions, awake flags and dynamic = motion type 2 pass
    through."""
    rng = np.random.default_rng(5)
    n = 400
    axis = rng.normal(size=(n, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    # away from 0 and 2 pi, where sqrt(1 - w^2) cancels (error ~ eps / sin^2(angle / 2))
    ang = rng.uniform(0.1, 2 * np.pi - 0.1, size=n)
    q = np.concatenate([axis * np.sin(ang / 2)[:, None], np.cos(ang / 2)[:, None]], axis=1)
    pos = rng.normal(size=(n, 3))
    awake = (rng.uniform(size=n) < 0.5).astype(np.uint8)
    mt = rng.integers(0, 3, size=n).astype(np.int32)
    op, aa, oa, dy = oracle.snapshot(pos, q, awake, mt)
    assert np.array_equal(op, pos) and np.array_equal(oa, awake.astype(bool)) and np.array_equal(dy, mt == 2)
    assert np.abs(aa[:, :3] - axis).max() < 1e-12
    assert np.abs(aa[:, 3] - ang).max() < 1e-12
    _, aa2, _, _ = oracle.snapshot(pos, q * rng.uniform(0.2, 5.0, size=(n, 1)), awake, mt)
    assert np.abs(aa2 - aa).max() < 1e-12
    special = np.array([[0, 0, 0, 1.0], [0, 0, 0, -1.0], [1e-12, 0, 0, 1.0], [0, 0, 1.0, 0.0],
                        [0, 1e-4, 0, 1.0]])
    _, sa, _, _ = oracle.snapshot(np.zeros((5, 3)), special, np.ones(5, np.uint8), np.full(5, 2, np.int32))
    assert np.array_equal(sa[:3], np.tile([0.0, 1.0, 0.0, 0.0], (3, 1)))
    assert np.allclose(sa[3], [0, 0, 1, np.pi], atol=1e-15)
    # a small rotation keeps its axis (sqrt(1 - w^2) carries ~1e-8 relative error here)
    assert np.allclose(sa[4], [0, 1, 0, 2 * np.arctan2(1e-4, 1.0)], rtol=1e-6, atol=0)


def _trigger_world(zone_shape, ball_radius, ball_pos, ball_vel):
    """The reference's trigger scenarios (tests/test_triggers.py): zero gravity, a
    static sensor at the origin, one dynamic ball."""
    from omi_physics_amd import PhysicsWorld, model
    w = PhysicsWorld(gravity=model.Gravity(gravity=0.0))
    zone = w.add_shape(zone_shape)
    w.add_body(model.Motion(type=model.STATIC), trigger=model.Trigger(shape=zone), position=(0, 0, 0))
    ball = w.add_shape(model.Shape.sphere(ball_radius))
    w.add_body(model.Motion(type=model.DYNAMIC, linearVelocity=ball_vel), collider=model.Collider(shape=ball),
               position=ball_pos)
    return w.export_arrays()


def test_trigger_events_on_the_reference_scenarios():
    """The oracle's trigger overlaps, diffed by TriggerSystem, on the reference's own
    trigger tests: a ball crossing a unit box zone fires enter before exit and flies
    straight through (velocity and y, z untouched); overlap holds exactly while the
    ball's centre is within r of the box (|x| < 0.5 + 0.2, away from the boundary);
    a ball sitting inside a zone gives enter then stays."""
    from omi_physics_amd import model
    from omi_physics_amd.triggers import TriggerSystem
    a = _trigger_world(model.Shape.box((1, 1, 1)), 0.2, (-3.0, 0, 0), (2.0, 0, 0))
    ow = oracle.OracleWorld(a)
    ts = TriggerSystem()
    events = []
    for k in range(120):
        x = ow.state()['position'][1, 0]           # the ball's centre when the step tests overlap
        ow.step(1.0 / 60.0)
        ov = ow.trigger_overlaps()
        if abs(abs(x) - 0.7) > 1e-6:
            assert (len(ov) == 1) == (abs(x) < 0.7), (k, x, ov)
        if len(ov):
            assert ov.tolist() == [[0, 1]]
        events += [(k, e) for e in ts.update(map(tuple, ov))]
    kinds = [e[0] for _, e in events]
    assert 'enter' in kinds and 'exit' in kinds and kinds.index('enter') < kinds.index('exit')
    assert kinds.count('enter') == 1 and kinds.count('exit') == 1
    st = ow.state()
    assert np.array_equal(st['linear_velocity'][1], [2.0, 0.0, 0.0])
    assert st['position'][1, 1] == 0.0 and st['position'][1, 2] == 0.0
    # a ball at rest inside a 2 m zone
    b = _trigger_world(model.Shape.box((2, 2, 2)), 0.2, (0, 0, 0), (0, 0, 0))
    ow = oracle.OracleWorld(b)
    ts = TriggerSystem()
    seen = []
    for _ in range(5):
        ow.step(1.0 / 60.0)
        seen += [e[0] for e in ts.update(map(tuple, ow.trigger_overlaps()))]
    assert seen[0] == 'enter' and seen.count('stay') >= 3


def test_trigger_box_zone_closed_form():
    """Axis-aligned unit box crossing a 2 x 2 x 2 box zone along x: the overlap holds
    exactly while the centres are closer than the half extents' sum (1.5) on x."""
    from omi_physics_amd import PhysicsWorld, model
    w = PhysicsWorld(gravity=model.Gravity(gravity=0.0))
    zone = w.add_shape(model.Shape.box((2, 2, 2)))
    w.add_body(model.Motion(type=model.STATIC), trigger=model.Trigger(shape=zone), position=(0, 0.2, -0.1))
    box = w.add_shape(model.Shape.box((1, 1, 1)))
    w.add_body(model.Motion(type=model.DYNAMIC, linearVelocity=(1.5, 0, 0)), collider=model.Collider(shape=box),
               position=(-3.0, 0, 0))
    ow = oracle.OracleWorld(w.export_arrays())
    hits = 0
    for k in range(240):
        x = ow.state()['position'][1, 0]
        ow.step(1.0 / 60.0)
        ov = ow.trigger_overlaps()
        if abs(abs(x) - 1.5) > 1e-6:
            assert (len(ov) == 1) == (abs(x) < 1.5), (k, x)
        hits += len(ov)
    assert hits > 100


def test_trigger_system_diff():
    from omi_physics_amd.triggers import TriggerSystem
    ts = TriggerSystem()
    assert ts.update([(3, 1), (0, 2)]) == [('enter', 0, 2), ('enter', 3, 1)]
    assert ts.update([(0, 2), (0, 5)]) == [('enter', 0, 5), ('stay', 0, 2), ('exit', 3, 1)]
    assert ts.update([]) == [('exit', 0, 2), ('exit', 0, 5)]
    assert ts.update([]) == [] and ts.overlaps == set()


# the reference's capsule-triangle geometry (tests/test_collide_batch.py)
CAP_FLOOR = ((-10.0, 0.0, -10.0), (10.0, 0.0, -10.0), (0.0, 0.0, 10.0))
CAP_WALL = ((1.0, -10.0, -10.0), (1.0, -10.0, 10.0), (1.0, 10.0, 0.0))


def _cap(centre, half=0.5, radius=0.4, rot=None):
    from omi_physics_amd.capsule import CapsuleProxy
    return CapsuleProxy(centre, half, radius, np.eye(3) if rot is None else rot)


def _ocap(cap, tris):
    return oracle.capsule_triangles(cap.p0, cap.p1, cap.radius, np.asarray(tris, dtype=float))


def test_capsule_triangles_hand_worked():
    """The oracle's capsule_triangles on the reference's hand-worked cases: resting
    0.2 into a floor (normal up, depth 0.2, point on the plane), clear of it, driven
    deep (depth 0.6 pushed down), wholly below, under a ceiling wound the other way,
    against a wall (depth 0.15), off a triangle's rim (push toward the corner), and a
    degenerate sliver (never a NaN normal)."""
    h, p, nrm, d = _ocap(_cap((0.0, 0.7, 0.0)), [CAP_FLOOR])
    assert h[0] and np.allclose(nrm[0], [0, 1, 0]) and d[0] == pytest.approx(0.2) and p[0][1] == 0.0
    assert not _ocap(_cap((0.0, 1.0, 0.0)), [CAP_FLOOR])[0][0]
    h, p, nrm, d = _ocap(_cap((0.0, -0.3, 0.0)), [CAP_FLOOR])
    assert h[0] and np.allclose(nrm[0], [0, -1, 0]) and d[0] == pytest.approx(0.6)
    assert not _ocap(_cap((0.0, -1.0, 0.0)), [CAP_FLOOR])[0][0]
    ceiling = ((-10.0, 0.0, -10.0), (0.0, 0.0, 10.0), (10.0, 0.0, -10.0))
    h, p, nrm, d = _ocap(_cap((0.0, -0.7, 0.0)), [ceiling])
    assert np.allclose(nrm[0], [0, -1, 0]) and d[0] == pytest.approx(0.2)
    h, p, nrm, d = _ocap(_cap((0.75, 0.0, 0.0)), [CAP_WALL])
    assert np.allclose(nrm[0], [-1, 0, 0]) and d[0] == pytest.approx(0.15)
    small = ((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    h, p, nrm, d = _ocap(_cap((-0.2, 0.0, -0.2), half=0.01, radius=0.4), [small])
    assert h[0] and np.allclose(p[0], [0, 0, 0]) and np.allclose(nrm[0], [-np.sqrt(0.5), 0, -np.sqrt(0.5)], atol=1e-6)
    sliver = ((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (2.0, 0.0, 0.0))
    h, p, nrm, d = _ocap(_cap((0.5, 0.0, 0.0), half=0.01, radius=0.4), [sliver])
    assert np.all(np.isfinite(nrm[0]))
    # no triangles; every triangle answered even when none touch
    assert len(_ocap(_cap((0, 0, 0)), np.zeros((0, 3, 3)))[0]) == 0
    h = _ocap(_cap((0.0, 50.0, 0.0)), [CAP_FLOOR, CAP_WALL])[0]
    assert len(h) == 2 and not h.any()


def _seg_tri_distance(p0, p1, tri, samples=400):
    """Brute force: the distance between a segment and a triangle, sampled densely."""
    a, b, c = (np.asarray(v, float) for v in tri)
    u = np.linspace(0.0, 1.0, samples)
    seg = p0[None, :] + u[:, None] * (p1 - p0)[None, :]
    g = np.linspace(0.0, 1.0, 120)
    s, t = np.meshgrid(g, g)
    keep = s + t <= 1.0
    pts = a + s[keep][:, None] * (b - a) + t[keep][:, None] * (c - a)
    best = np.inf
    for q in seg:
        best = min(best, np.sqrt(((pts - q) ** 2).sum(1)).min())
    return best


def test_capsule_triangles_against_brute_force():
    """Independent checks on random triangles: a face contact has a unit normal equal to
    +-the face normal and depth radius - height of a cap endpoint above the plane; a rim
    contact's depth is radius - the segment-triangle distance (brute force, 2e-2); no
    contact means the capsule is farther than its radius (brute force)."""
    rng = np.random.default_rng(7)
    cap = _cap((0.1, 0.2, -0.1), half=0.5, radius=0.7)
    tris = rng.normal(scale=1.2, size=(60, 3, 3))
    hit, pts, nrm, dep = _ocap(cap, tris)
    kinds = {'face': 0, 'rim': 0, 'none': 0}
    for i, tri in enumerate(tris):
        e = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        fnrm = e / np.linalg.norm(e)
        if not hit[i]:
            kinds['none'] += 1
            assert _seg_tri_distance(cap.p0, cap.p1, tri) > cap.radius - 2e-2, i
            continue
        assert abs(np.linalg.norm(nrm[i]) - 1.0) < 1e-12, i
        if abs(abs(np.dot(nrm[i], fnrm)) - 1.0) < 1e-12:
            kinds['face'] += 1
            heights = [np.dot(nrm[i], q - tri[0]) for q in (cap.p0, cap.p1)]
            assert min(abs(dep[i] - (cap.radius - h)) for h in heights) < 1e-12, i
        else:
            kinds['rim'] += 1
            assert abs(dep[i] - (cap.radius - _seg_tri_distance(cap.p0, cap.p1, tri))) < 2e-2, i
    assert min(kinds.values()) > 0, kinds


# ---- integrate_forces and the kinematic branches (backend.py:26-49,
#      world.py:381-384, solver.py:321-337), pinned by discrete closed forms ----

def _one_body(**motion):
    from omi_physics_amd import PhysicsWorld, model
    w = PhysicsWorld(gravity=model.Gravity(gravity=9.81, direction=(0, -1, 0)),
                     default_linear_damping=motion.pop('dlin', 0.0),
                     default_angular_damping=motion.pop('dang', 0.0), sleep_enabled=False)
    w.add_body(model.Motion(type=model.DYNAMIC, **motion), position=(0, 0, 0))
    return oracle.OracleWorld(w.export_arrays())


@pytest.mark.parametrize('c,c0', [(2.0, 0.0), (1.5, 0.5)])
def test_linear_damping_discrete_fixed_point(c, c0):
    """Reference test_physics_integrate.py:74-83 (|v| -> g/c) in its discrete form:
    v <- (v + g dt)(1 - (c + c0) dt) converges to g dt (1 - C dt) / (C dt), C = c + c0
    (per-body plus world default).  Damping applied before gravity would converge to
    g / C instead, 1 - C dt = 0.8 % away."""
    g, dt = 9.81, 1.0 / 240.0
    ow = _one_body(linearDamping=c, dlin=c0)
    ow.step(dt, 4000)
    C = c + c0
    expect = -g * dt * (1.0 - C * dt) / (C * dt)
    assert ow.state()['linear_velocity'][0, 1] == pytest.approx(expect, rel=1e-12)
    assert abs(ow.state()['linear_velocity'][0, 1]) == pytest.approx(g / C, rel=1e-2)


def test_quadratic_drag_discrete_fixed_point():
    """Reference test_physics_integrate.py:86-95 (|v| -> sqrt(g/k)) in its discrete form:
    with u = v + g dt and v <- u - k |u| u dt, the fixed point has k |u| u = g, so
    v* = -sqrt(g/k) + g dt along y; the drag uses the speed after gravity."""
    g, k, dt = 9.81, 0.5, 1.0 / 240.0
    ow = _one_body(quadraticDrag=k)
    ow.step(dt, 6000)
    expect = -np.sqrt(g / k) + g * dt
    assert ow.state()['linear_velocity'][0, 1] == pytest.approx(expect, rel=1e-12)


def test_damping_and_drag_together_fixed_point():
    """Both terms in the reference's order: gravity, then linear damping, then the drag
    on the damped speed.  Along y, falling, with f = 1 - C dt: u = v - g dt, w = f u,
    v' = w + k w^2 dt.  At the fixed point v = w / f + g dt, so w solves
    k dt w^2 + (1 - 1/f) w - g dt = 0 (its negative root)."""
    g, k, C, dt = 9.81, 0.3, 0.7, 1.0 / 120.0
    ow = _one_body(quadraticDrag=k, linearDamping=C)
    ow.step(dt, 8000)
    f = 1.0 - C * dt
    b = 1.0 - 1.0 / f
    w = (-b - np.sqrt(b * b + 4.0 * k * dt * g * dt)) / (2.0 * k * dt)
    assert ow.state()['linear_velocity'][0, 1] == pytest.approx(w / f + g * dt, rel=1e-11)


def test_angular_damping_geometric():
    """w_n = w_0 (1 - (c + c0) dt)^n for a free body (angular damping, world default added)."""
    dt, c, c0 = 1.0 / 60.0, 0.9, 0.3
    ow = _one_body(angularDamping=c, dang=c0, angularVelocity=(1.0, -2.0, 0.5), gravityFactor=0.0)
    ow.step(dt, 100)
    f = 1.0 - (c + c0) * dt
    np.testing.assert_allclose(ow.state()['angular_velocity'][0], np.array([1.0, -2.0, 0.5]) * f ** 100,
                               rtol=1e-12)


def test_damping_clamps_at_zero():
    """max(1 - c dt, 0): a damping larger than 1/dt stops the body outright each step."""
    ow = _one_body(linearDamping=100.0, angularDamping=100.0, angularVelocity=(3.0, 0, 0),
                   linearVelocity=(5.0, 0, 0))
    ow.step(1.0 / 60.0, 1)
    assert np.all(ow.state()['linear_velocity'][0] == 0.0) and np.all(ow.state()['angular_velocity'][0] == 0.0)


def test_kinematic_moves_at_constant_velocity_without_forces():
    """A kinematic body ignores gravity and damping and integrates x = x0 + n v dt;
    its orientation follows quat_integrate (a rotation about y by n w dt)."""
    from omi_physics_amd import PhysicsWorld, model
    w = PhysicsWorld(gravity=model.Gravity(gravity=9.81), default_linear_damping=0.5)
    w.add_body(model.Motion(type=model.KINEMATIC, linearVelocity=(1.5, -0.25, 0.5),
                            angularVelocity=(0.0, 0.6, 0.0), linearDamping=2.0), position=(1.0, 2.0, 3.0))
    ow = oracle.OracleWorld(w.export_arrays())
    dt, n = 1.0 / 60.0, 240
    ow.step(dt, n)
    S = ow.state()
    np.testing.assert_allclose(S['position'][0], np.array([1.0, 2.0, 3.0]) + n * dt * np.array([1.5, -0.25, 0.5]),
                               atol=1e-12, rtol=0)
    assert np.array_equal(S['linear_velocity'][0], [1.5, -0.25, 0.5])
    q = S['orientation'][0]
    # first-order quaternion steps renormalised: the angle is n atan(w dt / 2) * 2 exactly
    ang = 2.0 * n * np.arctan(0.5 * 0.6 * dt)
    np.testing.assert_allclose(q, [0.0, np.sin(ang / 2), 0.0, np.cos(ang / 2)], atol=1e-12)
    assert S['awake'][0] == 1


def test_kinematic_wakes_and_pushes_a_sleeper():
    """world.py:381-384 and solver.py:321-337: a sleeping box on the ground is no
    solid pair with the ground (neither is an awake mover), but a kinematic pusher
    sliding into it is an awake mover, wakes it on contact and pushes it along +x;
    a sleeper that nothing touches stays asleep with zero velocity."""
    from omi_physics_amd import PhysicsWorld, model
    w = PhysicsWorld(gravity=model.Gravity(gravity=9.81))
    g = w.add_shape(model.Shape.box((20, 1, 20)))
    b = w.add_shape(model.Shape.box((1, 1, 1)))
    w.add_body(model.Motion(type=model.STATIC), collider=model.Collider(shape=g), position=(0, -0.5, 0))
    w.add_body(model.Motion(), collider=model.Collider(shape=b), position=(0.0, 0.5, 0.0))
    w.add_body(model.Motion(), collider=model.Collider(shape=b), position=(6.0, 0.5, 6.0))
    w.add_body(model.Motion(type=model.KINEMATIC, linearVelocity=(1.0, 0, 0)),
               collider=model.Collider(shape=b), position=(-3.0, 0.5, 0.0))
    arrays = w.export_arrays()
    arrays['awake'] = np.array([1, 0, 0, 1], np.uint8)
    ow = oracle.OracleWorld(arrays)
    ow.step(1.0 / 60.0, 1)
    assert ow.counters()['solid_pairs'] == 0          # sleepers on the ground: no solid pair
    woke = None
    for s in range(2, 200):
        ow.step(1.0 / 60.0, 1)
        S = ow.state()
        if woke is None and S['awake'][1]:
            woke = s
            assert ow.counters()['contacts'] > 0      # woken by a contact with the pusher
    assert woke is not None and 100 <= woke <= 130      # the gap of 2 m closes at 1 m/s
    S = ow.state()
    assert S['position'][1, 0] > 0.5                   # pushed along +x
    assert S['awake'][2] == 0 and np.all(S['linear_velocity'][2] == 0)
    np.testing.assert_allclose(S['position'][3, 0], -3.0 + 199 / 60.0, atol=1e-12)


def test_update_sleep_rule():
    """ow_update_sleep against the island rule on a hand-built case: bodies 0-2 form an
    island with body 2 not yet ready, so none sleeps; bodies 3-4 form a ready island and
    both sleep (velocities zeroed); ready loner 5 sleeps; slow-but-young loner 6 keeps
    its timer running; fast body 7 wakes and resets; static 8 and kinematic 9 touch
    the first island and never sleep or count; timers add dt exactly."""
    n = 10
    mt = np.array([2, 2, 2, 2, 2, 2, 2, 2, 0, 1], np.int32)
    lv = np.zeros((n, 3)); av = np.zeros((n, 3))
    lv[:, 0] = 0.01; lv[7] = (0.5, 0, 0); av[4] = (0, 0.2, 0)
    awake = np.ones(n, np.uint8); awake[7] = 0
    timer = np.array([0.5, 0.45, 0.1, 0.9, 0.41, 0.5, 0.2, 3.0, 0.0, 0.0])
    ca = np.array([0, 1, 3, 0, 2], np.int32)
    cb = np.array([1, 2, 4, 8, 9], np.int32)
    dt = 1.0 / 60.0
    t0 = timer.copy()
    oracle.update_sleep(ca, cb, mt, lv, av, awake, timer, dt)
    assert list(awake) == [1, 1, 1, 0, 0, 0, 1, 1, 1, 1]
    slow = [0, 1, 2, 3, 4, 5, 6]
    assert np.array_equal(timer[slow], t0[slow] + dt) and timer[7] == 0.0
    assert not lv[3:6].any() and not av[3:6].any() and lv[0, 0] == 0.01
