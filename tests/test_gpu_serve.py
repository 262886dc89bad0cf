"""Continuous batching: admission, overlap, retirement; each world's result equals stepping it alone."""
import sys

import numpy as np
import pytest

from omi_physics_amd import scenes
from omi_physics_amd.batch import batch_from_arrays
from omi_physics_amd.serve import Request, RolloutServer, run_plain

pytestmark = pytest.mark.gpu


def make_requests(k, seed=1, steps=20):
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    for i in range(k):
        n = int(np.exp(rng.uniform(np.log(8), np.log(300))))
        arr = scenes.arrays(scenes.boxes_on_ground(n, seed=1000 + i, spread=4.0, y_range=(0.6, 6.0),
                                                   sphere_fraction=0.5))
        t += float(rng.exponential(0.004))
        reqs.append(Request(id=i, arrays=arr, steps=int(steps + rng.integers(0, 15)), arrival=t))
    return reqs


@pytest.mark.parametrize('precision', ['fp32', 'fp64'])
def test_served_results_equal_alone(precision):
    reqs = make_requests(24)
    print('serving 24 requests', precision, flush=True)
    # a small arena forces waiting for memory and slot reuse
    srv = RolloutServer(max_worlds=6, max_bodies=700, precision=precision)
    stats = srv.run(reqs)
    srv.close()
    assert len(srv.finished) == len(reqs)
    summ = stats.summary()
    assert summ['ticks'] < sum(r.steps for r in reqs)          # worlds overlapped on the device
    for r in reqs:
        alone = batch_from_arrays([r.arrays], precision=precision)
        alone.step(1.0 / 60.0, r.steps)
        alone.sync()
        ref = alone.read(0)
        alone.close()
        for key in ('position', 'orientation', 'linear_velocity', 'angular_velocity', 'awake'):
            assert np.array_equal(r.result[key], ref[key]), (r.id, key)


@pytest.mark.parametrize('precision', ['fp64'])
def test_unwaited_load_remove_fetch_keep_stream_order(precision):
    """Loads, removals, fetches and markers queued behind steps without waiting:
    a world fetched after its steps, its slot reloaded at once with another world,
    and a world that stays resident all equal their alone runs bit for bit."""
    def world(n, seed):
        return scenes.arrays(scenes.boxes_on_ground(n, seed=seed, spread=3.0, y_range=(0.6, 5.0),
                                                    sphere_fraction=0.5))
    w0, w1, w2 = world(40, 11), world(90, 12), world(120, 13)
    print('unwaited load/remove/fetch', precision, flush=True)

    def alone(w, steps):
        b = batch_from_arrays([w], precision=precision)
        b.step(1.0 / 60.0, steps)
        b.sync()
        r = b.read(0)
        b.close()
        return r
    from omi_physics_amd.batch import Batch
    b = Batch(max_worlds=2, max_bodies=400, precision=precision)
    b.load(0, w0)
    b.load(1, w1)
    b.step(1.0 / 60.0, 5)
    m = b.mark()
    t0 = b.fetch(0)
    b.remove(0)
    b.load(0, w2)                   # the same slot and body range, queued behind the fetch
    b.step(1.0 / 60.0, 3)
    t1 = b.fetch(0)
    b.mark_wait(m)
    r0, r2 = b.fetched(t0), b.fetched(t1)
    r1 = b.read(1)
    b.sync()
    b.close()
    for got, ref in ((r0, alone(w0, 5)), (r2, alone(w2, 3)), (r1, alone(w1, 8))):
        for key in ('position', 'orientation', 'linear_velocity', 'angular_velocity', 'awake', 'sleep_timer',
                    'aabb_min', 'aabb_max', 'prev_position', 'prev_orientation'):
            assert np.array_equal(got[key], ref[key]), key


def test_plain_form_runs():
    reqs = make_requests(3, seed=2, steps=5)
    out = run_plain(reqs)
    assert out['body_steps_per_s'] > 0 and all(r.result is not None for r in reqs)


def full_size_requests(k, rate, seed=5):
    """The serving workload's requests: Poisson arrivals of box/sphere worlds with
    log-uniform 8..4096 bodies, 60 steps each (scenes.serving_requests)."""
    spec = scenes.serving_requests(k, rate, seed=seed)
    return [Request(id=i, arrays=scenes.arrays(scenes.boxes_on_ground(r['n'], seed=r['seed'], sphere_fraction=0.5)),
                    steps=r['steps'], arrival=r['arrival']) for i, r in enumerate(spec)]


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_served_full_size_requests_equal_alone(precision):
    """28 requests of 8..4096 bodies x 60 steps arriving at 400/s into a server of
    12 slots and 16384 bodies (so requests wait for slots and memory and slots are
    reused): every result equals the request stepped alone, bit for bit; in FP64 the
    largest request and a mid-sized one equal the oracle."""
    import oracle
    reqs = full_size_requests(28, 400.0)
    sizes = [r.bodies for r in reqs]
    print('serving %d full-size requests %s: bodies %d..%d' % (len(reqs), precision, min(sizes), max(sizes)),
          flush=True)
    assert max(sizes) > 2000 and min(sizes) < 64
    srv = RolloutServer(max_worlds=12, max_bodies=16384, precision=precision)
    stats = srv.run(reqs)
    srv.close()
    assert len(srv.finished) == len(reqs)
    assert stats.summary()['ticks'] < sum(r.steps for r in reqs)
    for r in reqs:
        alone = batch_from_arrays([r.arrays], precision=precision)
        alone.step(1.0 / 60.0, r.steps)
        alone.sync()
        ref = alone.read(0)
        alone.close()
        for key in ('position', 'orientation', 'linear_velocity', 'angular_velocity', 'awake'):
            assert np.array_equal(r.result[key], ref[key]), (r.id, key)
    if precision == 'fp64':
        order = np.argsort(sizes)
        for i in (int(order[-1]), int(order[len(order) // 2])):
            ow = oracle.OracleWorld(reqs[i].arrays)
            ow.step(1.0 / 60.0, reqs[i].steps)
            for key in ('position', 'orientation', 'linear_velocity', 'angular_velocity', 'awake'):
                assert np.array_equal(reqs[i].result[key], ow.state()[key]), (i, key)
