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


def test_plain_form_runs():
    reqs = make_requests(3, seed=2, steps=5)
    out = run_plain(reqs)
    assert out['body_steps_per_s'] > 0 and all(r.result is not None for r in reqs)
