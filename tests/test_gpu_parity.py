"""HIP path vs the oracle on identical inputs (needs the GPU).

Bars: FP64 is bit-exact with the oracle (integer/order work and every float);
FP32 narrowphase/solver results are checked with the tolerances written in
each test.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from omi_physics_amd import _ffi, PhysicsWorld, model, scenes
from omi_physics_amd._ffi import _d, _i32, _i64, _u8, check, ptr, c64, ci32
from omi_physics_amd.batch import Batch, batch_from_arrays
from tests import helpers

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope='module')
def lib():
    L = _ffi.load()
    if L.omi_device_count() < 1:
        pytest.fail('no HIP device visible')
    return L


# ----------------------------------------------------------------- stages --

def gpu_refit(L, pos, quat, kind, half, margin=0.05):
    n = len(pos)
    pos, quat, half, kind = c64(pos), c64(quat), c64(half), ci32(kind)
    lo = np.empty((n, 3)); hi = np.empty((n, 3))
    z = np.zeros(n, np.int32); hp = np.zeros((1, 3))
    check(L.omi_refit(n, ptr(pos, _d), ptr(quat, _d), ptr(kind, _i32), ptr(half, _d), ptr(z, _i32),
                      ptr(z, _i32), ptr(hp, _d), 0, margin, ptr(lo, _d), ptr(hi, _d)))
    return lo, hi


def test_refit_bit_exact(lib, golden):
    g = golden('refit.npz')
    say('refit: 1200 bodies')
    lo, hi = gpu_refit(lib, g['pos'], g['quat'], g['kind'], g['half'])
    olo, ohi = oracle.refit(g['pos'], g['quat'], g['kind'], g['half'])
    assert np.array_equal(lo, olo) and np.array_equal(hi, ohi)
    assert np.array_equal(lo, g['lo']) and np.array_equal(hi, g['hi'])


def gpu_sap(L, lo, hi, mt, slots=None, tab=None):
    n = len(lo)
    lo, hi, mt = c64(lo), c64(hi), ci32(mt)
    slots = ci32(slots if slots is not None else np.zeros(n))
    tab = np.ascontiguousarray(tab if tab is not None else np.ones((1, 1)), np.uint8)
    cnt = np.zeros(1, np.int64)
    check(L.omi_sap_pairs(n, ptr(lo, _d), ptr(hi, _d), ptr(mt, _i32), ptr(slots, _i32), int(tab.shape[0]),
                          ptr(tab, _u8), None, 0, ptr(cnt, _i64)))
    out = np.empty((max(int(cnt[0]), 1), 2), np.int32)
    check(L.omi_sap_pairs(n, ptr(lo, _d), ptr(hi, _d), ptr(mt, _i32), ptr(slots, _i32), int(tab.shape[0]),
                          ptr(tab, _u8), ptr(out, _i32), int(cnt[0]), ptr(cnt, _i64)))
    return out[:int(cnt[0])]


@pytest.mark.parametrize('scene', range(4))
def test_sap_pairs_bit_exact(lib, golden, scene):
    g = golden('sap.npz')
    slots, tab = helpers.sap_filter_table(g['filt%d' % scene])
    say('sap scene', scene)
    pairs = gpu_sap(lib, g['lo%d' % scene], g['hi%d' % scene], g['mt%d' % scene], slots, tab)
    assert np.array_equal(pairs, g['pairs%d' % scene])


@pytest.mark.parametrize('n', [1, 2, 63, 4096, 4097, 9000])
def test_sap_pairs_sizes(lib, n):
    """Edge sizes: one body, two, the LDS/global-memory boundary, a large world; exact ties in lo.x."""
    rng = np.random.default_rng(n)
    c = rng.uniform(-20, 20, size=(n, 3)); c[:, 0] = np.round(c[:, 0] * 2) / 2
    h = rng.uniform(0.1, 0.8, size=(n, 3))
    lo, hi = c - h, c + h
    if n > 10:
        lo[0] = [-25, -1, -25]; hi[0] = [25, 1, 25]       # one ground-like box spanning the x range
    mt = rng.choice([0, 1, 2], size=n, p=[0.1, 0.1, 0.8]).astype(np.int32)
    say('sap size', n)
    ref, _ = oracle.sap_pairs(lo, hi, mt)
    assert np.array_equal(gpu_sap(lib, lo, hi, mt), ref)


def gpu_box_box(L, g, prec):
    n = len(g['ca'])
    args = [c64(g[k]) for k in ('ca', 'Ra', 'ha', 'cb', 'Rb', 'hb')]
    cnt = np.zeros(n, np.int32); nrm = np.zeros((n, 3)); pts = np.zeros((n, 4, 3)); dep = np.zeros((n, 4))
    check(L.omi_box_box(n, prec, *[ptr(a, _d) for a in args], ptr(cnt, _i32), ptr(nrm, _d), ptr(pts, _d),
                        ptr(dep, _d)))
    return cnt, nrm, pts, dep


def test_box_box_fp64_bit_exact(lib, golden):
    g = golden('boxbox.npz')
    say('box_box fp64: 2000 pairs')
    cnt, nrm, pts, dep = gpu_box_box(lib, g, 64)
    for i in range(len(cnt)):
        n, on, op, od = oracle.box_box(g['ca'][i], g['Ra'][i], g['ha'][i], g['cb'][i], g['Rb'][i], g['hb'][i])
        assert cnt[i] == n, i
        if n:
            assert np.array_equal(nrm[i], on) and np.array_equal(pts[i, :n], op) and np.array_equal(dep[i, :n], od), i


def test_box_box_fp32_close(lib, golden):
    """FP32 SAT can pick another axis or clip point on near-ties; 95% of manifolds must keep
    the FP64 count with normals within 1e-4 and points/depths within 1e-4 m."""
    g = golden('boxbox.npz')
    cnt, nrm, pts, dep = gpu_box_box(lib, g, 32)
    same = 0
    for i in range(len(cnt)):
        if cnt[i] != g['count'][i]:
            continue
        n = cnt[i]
        if n == 0 or (np.abs(nrm[i] - g['normal'][i]).max() < 1e-4 and np.abs(pts[i, :n] - g['points'][i][:n]).max() < 1e-4
                      and np.abs(dep[i, :n] - g['depths'][i][:n]).max() < 1e-4):
            same += 1
    assert same >= 0.95 * len(cnt)


def test_spheres_bit_exact(lib, golden):
    g = golden('spheres.npz')
    M = len(g['sa'])
    cnt = np.zeros(M, np.int32); nrm = np.zeros((M, 3)); pt = np.zeros((M, 3)); dep = np.zeros(M)
    check(lib.omi_sphere_sphere(M, 64, ptr(c64(g['sa']), _d), ptr(c64(g['ra']), _d), ptr(c64(g['sb']), _d),
                                ptr(c64(g['rb']), _d), ptr(cnt, _i32), ptr(nrm, _d), ptr(pt, _d), ptr(dep, _d)))
    for i in range(M):
        n, on, op, od = oracle.sphere_sphere(g['sa'][i], g['ra'][i], g['sb'][i], g['rb'][i])
        assert cnt[i] == n and (n == 0 or (np.array_equal(nrm[i], on) and np.array_equal(pt[i], op[0]) and dep[i] == od[0]))
    check(lib.omi_sphere_box(M, 64, ptr(c64(g['sc']), _d), ptr(c64(g['sr']), _d), ptr(c64(g['bc']), _d),
                             ptr(c64(g['bR']), _d), ptr(c64(g['bh']), _d), ptr(cnt, _i32), ptr(nrm, _d), ptr(pt, _d),
                             ptr(dep, _d)))
    for i in range(M):
        n, on, op, od = oracle.sphere_box(g['sc'][i], g['sr'][i], g['bc'][i], g['bR'][i], g['bh'][i])
        assert cnt[i] == n and (n == 0 or (np.array_equal(nrm[i], on) and np.array_equal(pt[i], op[0]) and dep[i] == od[0]))


def gpu_solve(L, g, isl, prec):
    pos = g['pos_in%d' % isl].copy(); lv = g['lv_in%d' % isl].copy(); av = g['av_in%d' % isl].copy()
    ni = g['seedn%d' % isl].copy(); ti = np.ascontiguousarray(g['seedt%d' % isl].copy())
    a = np.ascontiguousarray(g['a%d' % isl], np.int64); b = np.ascontiguousarray(g['b%d' % isl], np.int64)
    K, N = len(a), len(pos)
    check(L.omi_prepare_and_solve(K, N, ptr(a, _i64), ptr(b, _i64), ptr(c64(g['point%d' % isl]), _d),
                                  ptr(c64(g['normal%d' % isl]), _d), ptr(c64(g['depth%d' % isl]), _d), ptr(pos, _d),
                                  ptr(lv, _d), ptr(av, _d), ptr(c64(g['inv_mass_in%d' % isl]), _d),
                                  ptr(c64(g['invIw%d' % isl]), _d), ptr(c64(g['rest%d' % isl]), _d),
                                  ptr(c64(g['fric%d' % isl]), _d), ptr(ni, _d), ptr(ti, _d), 0.5, 10, 1, 0.8, 0.005,
                                  prec))
    return pos, lv, av, ni, ti


def test_prepare_and_solve_fp64_bit_exact(lib, golden):
    g = golden('solver.npz')
    for isl in range(12):
        say('prepare_and_solve fp64 island', isl)
        pos, lv, av, ni, ti = gpu_solve(lib, g, isl, 64)
        opos = g['pos_in%d' % isl].copy(); olv = g['lv_in%d' % isl].copy(); oav = g['av_in%d' % isl].copy()
        oni = g['seedn%d' % isl].copy(); oti = g['seedt%d' % isl].copy()
        oracle.prepare_and_solve(g['a%d' % isl], g['b%d' % isl], g['point%d' % isl], g['normal%d' % isl],
                                 g['depth%d' % isl], opos, olv, oav, g['inv_mass_in%d' % isl], g['invIw%d' % isl],
                                 g['rest%d' % isl], g['fric%d' % isl], oni, oti)
        for x, y in ((pos, opos), (lv, olv), (av, oav), (ni, oni), (ti, oti)):
            assert np.array_equal(x, y)


def test_prepare_and_solve_fp32_close(lib, golden):
    """FP32 solver rows: velocities within 1e-4 m/s·scale, impulses within 1e-4 relative to the largest."""
    g = golden('solver.npz')
    for isl in range(12):
        pos, lv, av, ni, ti = gpu_solve(lib, g, isl, 32)
        scale = max(1.0, np.abs(g['lv_out%d' % isl]).max())
        assert np.abs(lv - g['lv_out%d' % isl]).max() < 2e-4 * scale
        assert np.abs(pos - g['pos_out%d' % isl]).max() < 1e-5
        assert np.abs(ni - g['nimp_out%d' % isl]).max() < 2e-4 * max(1.0, np.abs(g['nimp_out%d' % isl]).max())


# ------------------------------------------------------------ whole steps --

def run_both(arrays, steps, precision='fp64', check_every=10):
    """Step the same world on the GPU and in the oracle; yield (step, gpu, oracle) states."""
    b = batch_from_arrays([arrays], precision=precision, record_broadphase=True)
    ow = oracle.OracleWorld(arrays)
    done = 0
    while done < steps:
        k = min(check_every, steps - done)
        b.step(DT, k)
        b.sync()
        ow.step(DT, k)
        done += k
        yield done, b, ow


def assert_state_equal(b, ow, where):
    g = b.read(0); o = ow.state()
    for key in ('position', 'orientation', 'linear_velocity', 'angular_velocity', 'aabb_min', 'aabb_max',
                'sleep_timer'):
        if not np.array_equal(g[key], o[key]):
            raise AssertionError('%s: %s differs (max %.3g)' % (where, key, np.abs(g[key] - o[key]).max()))
    assert np.array_equal(g['awake'], o['awake']), where


@pytest.mark.parametrize('name', ['dropbox', 'stack5', 'rolling', 'mixed24'])
def test_small_scenes_bit_exact(lib, name):
    arrays = helpers.world(helpers.traj_spec(name)).export_arrays()
    say('scene', name)
    for s, b, ow in run_both(arrays, 120):
        assert_state_equal(b, ow, '%s step %d' % (name, s))


def test_pile_bit_exact_with_pairs_and_contacts(lib):
    arrays = helpers.world(helpers.pile(160, seed=5)).export_arrays()
    say('pile 161 bodies, 150 steps fp64')
    seen_contacts = 0
    for s, b, ow in run_both(arrays, 150, check_every=5):
        assert_state_equal(b, ow, 'pile step %d' % s)
        assert np.array_equal(b.pairs(0, solid=False), ow.pairs(False)), s
        assert np.array_equal(b.pairs(0, solid=True), ow.pairs(True)), s
        gc, oc = b.contacts(0), ow.contacts()
        for key in ('a', 'b', 'point', 'normal', 'depth', 'normal_impulse', 'tangent_impulse'):
            assert np.array_equal(gc[key], oc[key]), (s, key)
        seen_contacts = max(seen_contacts, len(gc['a']))
    assert seen_contacts > 100


def test_boxes_on_ground_1000_bit_exact(lib):
    """The single-world FP64 workload (1000 boxes, 1 ground), 300 steps."""
    arrays = helpers.world(scenes.boxes_on_ground(1000, seed=1)).export_arrays()
    say('1000 boxes fp64 300 steps')
    for s, b, ow in run_both(arrays, 300, check_every=50):
        say('  step', s, b.counters())
        assert_state_equal(b, ow, 'boxes step %d' % s)


def test_batch_equals_alone(lib):
    """Each world's result is the same alone and packed with others (bit-exact)."""
    specs = [helpers.pile(n, seed=s) for n, s in ((40, 1), (90, 2), (15, 3), (130, 4))]
    arrays = [helpers.world(sp).export_arrays() for sp in specs]
    say('batch of 4 vs alone')
    b = batch_from_arrays(arrays)
    b.step(DT, 60); b.sync()
    for i, a in enumerate(arrays):
        alone = batch_from_arrays([a]); alone.step(DT, 60); alone.sync()
        x, y = b.read(i), alone.read(0)
        for key in ('position', 'orientation', 'linear_velocity', 'angular_velocity', 'awake'):
            assert np.array_equal(x[key], y[key]), (i, key)


def test_full_4096_world_batch_fp64_matches_oracle_sample(lib):
    """The batched workload at its full size: 4096 box/sphere worlds (lognormal
    sizes, 436778 bodies) stepped 300 times in FP64 in one graph-captured batch,
    with no capacity overflow; the three largest worlds and five spread over the
    slots equal the oracle bit for bit."""
    sizes = scenes.world_sizes(4096, seed=3)
    arrays = [scenes.arrays(scenes.boxes_on_ground(int(sizes[w]), seed=3 * 100003 + w, sphere_fraction=0.5))
              for w in range(4096)]
    say('4096 worlds, %d bodies, fp64 300 steps' % sum(a['n'] for a in arrays))
    b = batch_from_arrays(arrays, precision='fp64')
    b.step(DT, 300)
    b.sync()
    assert b.counters()['overflow'] == 0
    largest = [int(i) for i in np.argsort([-a['n'] for a in arrays], kind='stable')[:3]]
    for i in largest + [0, 5, 100, 2000, 4095]:
        ow = oracle.OracleWorld(arrays[i])
        ow.step(DT, 300)
        g, o = b.read(i), ow.state()
        for key in ('position', 'orientation', 'linear_velocity', 'angular_velocity', 'awake'):
            assert np.array_equal(g[key], o[key]), (i, key)
    b.close()


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_group_lanes_equal_single_batch(lib, precision):
    """Worlds of three size classes split into lanes (one member batch per class,
    stepped side by side) end bit-identical to one batch holding them all: states,
    contacts, per-world statistics and counters (schedule depths aside), in graph
    mode, in eager timed mode and in graph mode again after it."""
    sizes = ((40, 1), (600, 2), (1500, 3), (130, 4))
    arrays = [scenes.arrays(scenes.boxes_on_ground(n, seed=s, spread=10.0, y_range=(0.6, 8.0),
                                                   sphere_fraction=0.5)) for n, s in sizes]
    say('group of lanes vs one batch', precision, [a['n'] for a in arrays])
    grp = batch_from_arrays(arrays, precision=precision, lanes=True)
    one = batch_from_arrays(arrays, precision=precision)
    assert hasattr(grp, 'lanes') and len(grp.lanes) == 3 and not hasattr(one, 'lanes')
    grp.prepare(DT)
    for phase, timed in enumerate((False, True, False)):
        grp.set_timing(timed); one.set_timing(timed)
        grp.step(DT, 15); one.step(DT, 15)
        grp.sync(); one.sync()
        say('  phase', phase, 'timed' if timed else 'graph', grp.counters())
        for i in range(len(arrays)):
            x, y = grp.read(i), one.read(i)
            for key in ('position', 'orientation', 'linear_velocity', 'angular_velocity', 'awake'):
                assert np.array_equal(x[key], y[key]), (phase, i, key)
            cx, cy = grp.contacts(i), one.contacts(i)
            for key in ('a', 'b', 'normal', 'depth', 'normal_impulse'):
                assert np.array_equal(cx[key], cy[key]), (phase, i, key)
        # (schedule depths may differ: a class alone in its lane can be solved
        # island by island, and a world's long schedule may shift its last
        # passes; every schedule gives the same results)
        sg, so = grp.world_stats(), one.world_stats()
        for key in ('bodies', 'contacts', 'manifolds', 'size_class'):
            assert np.array_equal(sg[key], so[key]), (phase, key)
        assert (sg['levels'] > 0).sum() == (so['levels'] > 0).sum(), phase
        cg, co = grp.counters(), one.counters()
        for key in ('bodies', 'x_candidates', 'broadphase_pairs', 'solid_pairs', 'manifolds', 'contacts'):
            assert cg[key] == co[key], (phase, key)
        if timed:
            lt = grp.lane_kernel_times()
            assert len(lt) == 3 and all(t['solve_top'] > 0 for t in lt)
    assert grp.counters()['contacts'] > 0
    for x, y in zip(grp.snapshot(), one.snapshot()):
        for u, v in zip(x, y):
            assert np.array_equal(u, v)
    grp.close(); one.close()


def test_snapshot_matches_oracle(lib):
    """Pose snapshots packed on the device (threaded.py:149-158 format) against the
    oracle's restatement on the same state.  Positions, awake and dynamic flags are
    exact; the axis, a quotient of correctly rounded operations, within 1e-15; the
    angle 2 acos(w) within 1e-14 rad (the device's acos against the host's, each
    within an ulp or two).  A subset in another order gives the same rows."""
    arrays = [helpers.world(helpers.pile(n, seed=s)).export_arrays() for n, s in ((40, 1), (90, 2), (15, 3))]
    say('snapshot of 3 worlds after 40 steps')
    b = batch_from_arrays(arrays)
    b.step(DT, 40); b.sync()
    snaps = b.snapshot()
    sub = b.snapshot([2, 0])
    assert len(snaps) == 3 and len(sub) == 2
    for i, (p, aa, aw, dy) in enumerate(snaps):
        st = b.read(i)
        op, oaa, oaw, ody = oracle.snapshot(st['position'], st['orientation'], st['awake'],
                                            arrays[i]['motion_type'])
        assert p.shape == (arrays[i]['n'], 3) and aa.shape == (arrays[i]['n'], 4)
        assert np.array_equal(p, op) and np.array_equal(aw, oaw) and np.array_equal(dy, ody)
        assert np.abs(aa[:, :3] - oaa[:, :3]).max() <= 1e-15, i
        assert np.abs(aa[:, 3] - oaa[:, 3]).max() <= 1e-14, i
        assert dy.sum() == arrays[i]['n'] - 1                 # every body but the ground
    for (k, i) in ((0, 2), (1, 0)):
        for u, v in zip(sub[k], snaps[i]):
            assert np.array_equal(u, v)
    b.close()


@pytest.mark.parametrize('kind', ['max_contacts', 'max_pairs'])
def test_capacity_overflow_drops_whole_worlds(lib, kind):
    """A step that runs out of contact (or pair) capacity raises on sync; every world
    either keeps all of its contacts, with a state bit-equal to an unconstrained batch,
    or gets none."""
    specs = [helpers.pile(n, seed=s) for n, s in ((40, 1), (90, 2), (15, 3), (130, 4), (70, 5), (110, 6))]
    arrays = [helpers.world(sp).export_arrays() for sp in specs]
    W = len(arrays)
    probe = batch_from_arrays(arrays)
    peak = 0
    for _ in range(12):
        probe.step(DT, 10); probe.sync()
        peak = max(peak, int(probe.counters()['contacts' if kind == 'max_contacts' else 'solid_pairs']))
    cap = int(peak * 0.7)
    say('overflow of %s: peak %d, capacity %d' % (kind, peak, cap))
    big = batch_from_arrays(arrays)
    small = batch_from_arrays(arrays, **{kind: cap})
    for s in range(120):
        big.step(DT); big.sync()
        small.step(DT)
        try:
            small.sync()
            over = False
        except RuntimeError as e:
            assert 'capacity' in str(e)
            over = True
        kb, ks = big.world_stats(W)['contacts'], small.world_stats(W)['contacts']
        if not over:
            assert np.array_equal(kb, ks), s
            continue
        fit = ks > 0
        assert np.all((ks == kb) | (ks == 0)), (s, kb, ks)
        assert (kb[~fit] > 0).any(), (s, kb, ks)
        if kind == 'max_contacts':
            assert ks.sum() <= cap
        for i in np.flatnonzero(fit):
            for key in ('position', 'orientation', 'linear_velocity', 'angular_velocity'):
                assert np.array_equal(small.read(i)[key], big.read(i)[key]), (s, i, key)
        return
    pytest.fail('the capacity was never exceeded')


def test_world_reload_resets_slot(lib):
    """Removing a world and loading another into the freed slot gives the fresh result."""
    a = helpers.world(helpers.pile(60, seed=8)).export_arrays()
    c = helpers.world(helpers.pile(50, seed=9)).export_arrays()
    b = Batch(2, 200)
    b.load(0, a); b.load(1, c)
    b.step(DT, 30); b.sync()
    b.remove(0)
    b.load(0, c)
    b.step(DT, 20); b.sync()
    fresh = batch_from_arrays([c]); fresh.step(DT, 20); fresh.sync()
    assert np.array_equal(b.read(0)['position'], fresh.read(0)['position'])


def test_physicsworld_api_on_hip(lib):
    """The reference-shaped API: add_body, step, column views; free fall closed form."""
    w = PhysicsWorld(gravity=model.Gravity(gravity=9.81, direction=(0, -1, 0)))
    body = w.add_body(model.Motion(type=model.DYNAMIC), position=(0, 0, 0))
    dt, n = 1.0 / 240.0, 240
    for _ in range(n):
        w.step(dt)
    assert w.position[body][1] == pytest.approx(-0.5 * 9.81 * dt * dt * n * (n + 1), rel=1e-9)


def test_physicsworld_resting_sphere(lib):
    """Reference test_solver.py:31-37: a dropped sphere comes to rest at its radius."""
    w = PhysicsWorld(gravity=model.Gravity(gravity=9.81), sleep_enabled=False)
    mat = w.add_material(model.Material(staticFriction=0.6, dynamicFriction=0.6))
    g = w.add_shape(model.Shape.box((40, 1, 40)))
    w.add_body(model.Motion(type=model.STATIC), collider=model.Collider(shape=g, physicsMaterial=mat),
               position=(0, -0.5, 0))
    s = w.add_shape(model.Shape.sphere(0.5))
    body = w.add_body(model.Motion(mass=1.0), collider=model.Collider(shape=s, physicsMaterial=mat),
                      position=(0, 3.0, 0))
    for _ in range(600):
        w.step(1.0 / 120.0)
    assert w.position[body][1] == pytest.approx(0.5, abs=0.05)
    assert len(w.contacts) == 1 and w.contacts[0].normal_impulse > 0


def test_fp32_trajectory_error(lib):
    """Mixed precision (FP32 narrowphase + solver, FP64 integration) against the FP64 oracle:
    free flight stays exact to 1e-9 m; the pile stays within 5e-3 m over 60 steps."""
    arrays = helpers.world(helpers.pile(100, seed=12)).export_arrays()
    b = batch_from_arrays([arrays], precision='fp32')
    ow = oracle.OracleWorld(arrays)
    b.step(DT, 60); b.sync(); ow.step(DT, 60)
    err = np.abs(b.read(0)['position'] - ow.state()['position']).max()
    say('fp32 pile 60 steps max err', err)
    assert err < 5e-3
    spec = scenes.new_spec()
    scenes._ground(spec)
    spec['shapes'].append(('box', (1.0, 1.0, 1.0)))
    for i in range(5):
        for j in range(5):
            for k in range(4):      # 2.5 m apart, 30-40 m up: no contact within 60 steps
                spec['bodies'].append(dict(type='dynamic', shape=1, mass=1.0,
                                           position=(2.5 * i - 5, 30.0 + 2.5 * k, 2.5 * j - 5),
                                           orientation=(0.0, 0.0, 0.0, 1.0), material=0,
                                           angular_velocity=(0.3 * i, 0.1 * j, -0.2 * k)))
    fall = helpers.world(spec).export_arrays()
    b = batch_from_arrays([fall], precision='fp32')
    ow = oracle.OracleWorld(fall)
    b.step(DT, 60); b.sync(); ow.step(DT, 60)
    assert b.counters()['contacts'] == 0
    # free flight never touches FP32: positions and orientations are the FP64 ones
    assert np.array_equal(b.read(0)['position'], ow.state()['position'])
    assert np.array_equal(b.read(0)['orientation'], ow.state()['orientation'])


def test_collide_convex_bit_exact(lib, golden):
    g = golden('gjk.npz')
    hp, ho = g['hull_pts'], g['hull_off']
    G = len(g['idx'])
    ka = np.full(G, 3, np.int32); kb = np.where(np.arange(G) % 10 == 9, 2, 3).astype(np.int32)
    oa = ho[g['idx'][:, 0]].astype(np.int32); na = (ho[g['idx'][:, 0] + 1] - ho[g['idx'][:, 0]]).astype(np.int32)
    ob = ho[g['idx'][:, 1]].astype(np.int32); nb = (ho[g['idx'][:, 1] + 1] - ho[g['idx'][:, 1]]).astype(np.int32)
    hb = np.tile([0.8, 0.3, 0.6], (G, 1)); ha = np.zeros((G, 3))
    ca, cb = c64(g['pos'][:, 0]), c64(g['pos'][:, 1]); Ra, Rb = c64(g['R'][:, 0]), c64(g['R'][:, 1])
    cnt = np.zeros(G, np.int32); nrm = np.zeros((G, 3)); pt = np.zeros((G, 3)); dep = np.zeros(G)
    say('collide_convex: 300 hull pairs')
    check(lib.omi_collide_convex(G, 64, ptr(ka, _i32), ptr(ca, _d), ptr(Ra, _d), ptr(c64(ha), _d), ptr(oa, _i32),
                                 ptr(na, _i32), ptr(kb, _i32), ptr(cb, _d), ptr(Rb, _d), ptr(c64(hb), _d),
                                 ptr(ob, _i32), ptr(nb, _i32), ptr(c64(hp), _d), len(hp), ptr(cnt, _i32),
                                 ptr(nrm, _d), ptr(pt, _d), ptr(dep, _d)))
    for k in range(G):
        i, j = g['idx'][k]
        n, on, op, od = oracle.collide_convex(3, ca[k], Ra[k], None, hp[ho[i]:ho[i + 1]], int(kb[k]), cb[k], Rb[k],
                                              hb[k] if kb[k] == 2 else None, hp[ho[j]:ho[j + 1]])
        assert cnt[k] == n, k
        if n:
            assert np.array_equal(nrm[k], on) and np.array_equal(pt[k], op) and dep[k] == od, k


def test_convex_world_bit_exact(lib):
    """Cooked hulls falling on the ground: refit from hull points and GJK/EPA contacts in the step."""
    arrays = helpers.world(scenes.convex_heap(60, seed=6, spread=2.0, y_range=(0.8, 5.0))).export_arrays()
    say('convex heap 61 bodies, 90 steps fp64')
    contacts = 0
    for s, b, ow in run_both(arrays, 90, check_every=10):
        assert_state_equal(b, ow, 'convex step %d' % s)
        contacts = max(contacts, b.counters()['contacts'])
    assert contacts > 10


def test_pyramids_and_debris_world_fp64_bit_exact(lib):
    """The large single world at its full size (20 ten-level pyramids + 8000
    box/sphere debris, 9101 bodies; its solver state lives in HBM): 40 FP64 steps
    equal the oracle bit for bit."""
    arrays = scenes.arrays(scenes.pyramids_and_debris())
    say('pyramids and debris, 9101 bodies, fp64 40 steps')
    for s, b, ow in run_both(arrays, 40, check_every=20):
        assert_state_equal(b, ow, 'pyramids step %d' % s)


def test_convex_and_chains_world_fp64_bit_exact(lib):
    """Hulls and hinge chains in one world (the convex-and-joints workload at
    300 hulls and 6 chains of 30 links): GJK/EPA contacts, joints in LDS and the
    contact solve together, 200 FP64 steps equal the oracle bit for bit."""
    arrays = scenes.build(scenes.convex_and_chains(n_hulls=300, n_chains=6, links=30, seed=5),
                          PhysicsWorld, model).export_arrays()
    say('300 hulls + 6 chains x 30, fp64 200 steps')
    contacts = 0
    for s, b, ow in run_both(arrays, 200, check_every=50):
        assert_state_equal(b, ow, 'convex+chains step %d' % s)
        contacts = max(contacts, b.counters()['contacts'])
    assert contacts > 100


def test_joint_chains_bit_exact(lib):
    """PointConstraint chains (joints.py:60-121) with contacts against the ground, 120 steps FP64."""
    arrays = helpers.world(scenes.hinge_chains(n_chains=3, links=8, seed=5, spacing=1.5,
                                               top=10.0)).export_arrays()
    say('3 chains x 8 links, fp64 120 steps')
    assert len(arrays['joint_a']) == 24
    for s, b, ow in run_both(arrays, 120, check_every=20):
        assert_state_equal(b, ow, 'chains step %d' % s)


def test_joint_sweeps_in_lds_and_in_body_arrays(lib):
    """60 chains of 50 links touch more joint bodies (3000) than the joint sweeps
    stage in LDS, so that world's sweeps update the body arrays; the small chain
    world in the same batch sweeps in LDS.  Both FP64 bit-exact for 40 steps."""
    big = helpers.world(scenes.hinge_chains(n_chains=60, links=50, seed=7, top=30.0)).export_arrays()
    small = helpers.world(scenes.hinge_chains(n_chains=2, links=6, seed=8, spacing=1.5, top=4.0)).export_arrays()
    assert len(big['joint_a']) == 3000
    say('chains 3000 + 12 links, fp64 40 steps')
    b = batch_from_arrays([big, small], precision='fp64')
    ows = [oracle.OracleWorld(big), oracle.OracleWorld(small)]
    for _ in range(4):
        b.step(DT, 10)
        b.sync()
        for slot, ow in enumerate(ows):
            ow.step(DT, 10)
            g, o = b.read(slot), ow.state()
            for key in ('position', 'orientation', 'linear_velocity', 'angular_velocity'):
                assert np.array_equal(g[key], o[key]), (slot, key, np.abs(g[key] - o[key]).max())
    assert np.abs(ows[0].state()['angular_velocity']).max() > 0.1


def test_joint_schedule_rebuilt_on_reload(lib):
    """The joint level schedule and joint-body LDS records are kept across steps
    and rebuilt when a slot is reloaded: a small chain world steps, its slot is
    refilled with a larger chain world (more joint bodies than the batch had
    planned LDS for), and a second world joins; all stay FP64 bit-exact."""
    first = helpers.world(scenes.hinge_chains(n_chains=2, links=5, seed=9, spacing=1.5, top=6.0)).export_arrays()
    second = helpers.world(scenes.hinge_chains(n_chains=6, links=40, seed=10, spacing=1.5, top=30.0)).export_arrays()
    third = helpers.world(scenes.hinge_chains(n_chains=3, links=8, seed=11, spacing=1.5, top=10.0)).export_arrays()
    n = max(first['n'], second['n']) + third['n']
    b = Batch(2, n, precision='fp64', max_joints=len(second['joint_a']) + len(third['joint_a']))
    say('joint schedule: 10 links, then 240 links in the same slot, then a third world')
    b.load(0, first)
    ow = oracle.OracleWorld(first)
    b.step(DT, 20); b.sync(); ow.step(DT, 20)
    assert np.array_equal(b.read(0)['position'], ow.state()['position'])
    b.remove(0)
    b.load(0, second)
    b.load(1, third)
    ows = [oracle.OracleWorld(second), oracle.OracleWorld(third)]
    for _ in range(3):
        b.step(DT, 10); b.sync()
        for slot, o in enumerate(ows):
            o.step(DT, 10)
            g, st = b.read(slot), o.state()
            for key in ('position', 'orientation', 'linear_velocity', 'angular_velocity'):
                assert np.array_equal(g[key], st[key]), (slot, key, np.abs(g[key] - st[key]).max())
    b.close()


def test_physicsworld_pendulum_on_hip(lib):
    """A ball joint to the world keeps a pendulum's bob at its rod length (joints.py:60-121)."""
    from omi_physics_amd.joints import PointConstraint
    w = PhysicsWorld(gravity=model.Gravity(gravity=9.81), sleep_enabled=False)
    s = w.add_shape(model.Shape.sphere(0.1))
    bob = w.add_body(model.Motion(mass=1.0), collider=model.Collider(shape=s), position=(1.0, 2.0, 0.0))
    w.add_joint_constraint(PointConstraint(-1, bob, (0.0, 2.0, 0.0)))
    ow = oracle.OracleWorld(w.export_arrays())
    for _ in range(240):
        w.step(1.0 / 120.0)
        r = np.linalg.norm(np.asarray(w.position[bob]) - np.array([0.0, 2.0, 0.0]))
        assert r == pytest.approx(1.0, abs=0.02)
    ow.step(1.0 / 120.0, 240)
    assert np.array_equal(np.asarray(w.position), ow.state()['position'])
    assert w.position[bob][1] < 2.0


def test_trigger_overlaps_match_oracle(lib):
    """Sensors on the device against the oracle, step by step: a pile falling through
    static box, sphere and rotated-box trigger zones and past a dynamic sensor sphere.
    The overlap sets (trigger body, other body) are identical every step, and the bodies'
    states stay bit-identical (a sensor gives no impulse)."""
    arrays = helpers.world(helpers.trigger_scene()).export_arrays()
    say('trigger scene: %d bodies' % arrays['n'])
    b = batch_from_arrays([arrays], precision='fp64')
    ow = oracle.OracleWorld(arrays)
    total = 0
    kinds = set()
    for s in range(90):
        b.step(DT, 1); b.sync()
        ow.step(DT, 1)
        g, o = b.trigger_overlaps(0), ow.trigger_overlaps()
        assert np.array_equal(g, o), (s, g.tolist(), o.tolist())
        total += len(g)
        kinds |= {int(t) for t in g[:, 0]}
        if s % 10 == 9:
            assert_state_equal(b, ow, 'trigger scene step %d' % s)
    say('  overlaps over 90 steps:', total, 'sensors seen:', sorted(kinds))
    assert total > 50 and len(kinds) >= 3
    b.close()


def test_physicsworld_trigger_events_on_hip(lib):
    """The reference's trigger tests (tests/test_triggers.py) through PhysicsWorld on the
    device: a ball crossing a box zone fires enter before exit and keeps its velocity
    and line; a ball resting inside a zone gets enter, then stays."""
    def scene(zone_size, r, pos, vel):
        w = PhysicsWorld(gravity=model.Gravity(gravity=0.0))
        zone = w.add_shape(model.Shape.box(zone_size))
        w.add_body(model.Motion(type=model.STATIC), trigger=model.Trigger(shape=zone), position=(0, 0, 0))
        s = w.add_shape(model.Shape.sphere(r))
        ball = w.add_body(model.Motion(type=model.DYNAMIC, linearVelocity=vel), collider=model.Collider(shape=s),
                          position=pos)
        return w, ball
    w, ball = scene((1, 1, 1), 0.2, (-3.0, 0, 0), (2.0, 0, 0))
    events = []
    w.add_trigger_listener(lambda kind, t, o: events.append((kind, t, o)))
    for _ in range(120):
        w.step(DT)
    kinds = [e[0] for e in events]
    say('crossing:', kinds.count('enter'), 'enter', kinds.count('stay'), 'stay', kinds.count('exit'), 'exit')
    assert 'enter' in kinds and 'exit' in kinds and kinds.index('enter') < kinds.index('exit')
    assert all(t == 0 and o == ball for _, t, o in events)
    assert np.allclose(w.linear_velocity[ball], (2.0, 0, 0))
    assert w.position[ball][1] == pytest.approx(0.0, abs=1e-9)
    assert w.position[ball][2] == pytest.approx(0.0, abs=1e-9)
    w2, _ = scene((2, 2, 2), 0.2, (0, 0, 0), (0, 0, 0))
    seen = []
    w2.add_trigger_listener(lambda kind, t, o: seen.append(kind))
    for _ in range(5):
        w2.step(DT)
    assert seen[0] == 'enter' and seen.count('stay') >= 3


def test_capsule_triangles_bit_exact(lib):
    """omi_capsule_triangles against the oracle's restatement of the Cython loop, bit for
    bit: 5000 random triangles around an upright and a tilted capsule, the reference's
    hand-worked floor / wall / ceiling / rim / sliver cases, and an empty batch."""
    from omi_physics_amd.capsule import CapsuleProxy, capsule_triangle_batch, capsule_triangles
    rng = np.random.default_rng(3)
    tris = rng.normal(scale=1.2, size=(5000, 3, 3))
    a = 0.6
    tilt = np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]])
    caps = [CapsuleProxy((0.1, 0.2, -0.1), 0.5, 0.7, np.eye(3)), CapsuleProxy((0.0, 0.3, 0.0), 0.5, 0.4, tilt)]
    hand = np.array([((-10.0, 0.0, -10.0), (10.0, 0.0, -10.0), (0.0, 0.0, 10.0)),
                     ((1.0, -10.0, -10.0), (1.0, -10.0, 10.0), (1.0, 10.0, 0.0)),
                     ((-10.0, 0.0, -10.0), (0.0, 0.0, 10.0), (10.0, 0.0, -10.0)),
                     ((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0)),
                     ((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (2.0, 0.0, 0.0))])
    for cap in caps + [CapsuleProxy(c, h, r, np.eye(3)) for c, h, r in
                       (((0.0, 0.7, 0.0), 0.5, 0.4), ((0.0, -0.3, 0.0), 0.5, 0.4), ((0.75, 0.0, 0.0), 0.5, 0.4),
                        ((-0.2, 0.0, -0.2), 0.01, 0.4), ((0.5, 0.0, 0.0), 0.01, 0.4))]:
        for batch in (tris, hand):
            g = capsule_triangle_batch(cap, batch)
            o = oracle.capsule_triangles(cap.p0, cap.p1, cap.radius, batch)
            assert np.array_equal(g[0], o[0])
            for x, y in zip(g[1:], o[1:]):
                assert np.array_equal(x, y)
    h, p, n, d = capsule_triangles(caps[0], tris[:, 0], tris[:, 1], tris[:, 2])
    say('capsule: %d of 5000 random triangles hit' % int(h.sum()))
    assert np.array_equal(h, oracle.capsule_triangles(caps[0].p0, caps[0].p1, 0.7, tris)[0])
    empty = capsule_triangle_batch(caps[0], np.zeros((0, 3, 3)))
    assert all(len(x) == 0 for x in empty)
