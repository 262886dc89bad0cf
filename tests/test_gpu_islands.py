"""Island-by-island solves (SequentialImpulseSolver.build_islands,
solver.py:412-445).

A size class that holds too few worlds to fill the GPU is solved island by
island: each island (bodies joined by contacts whose two sides are dynamic)
sweeps its contacts in their order in the world, so the results equal those
of the one-block-per-world solve bit for bit, in FP64 and in FP32 alike, and
FP64 equals the oracle.
"""
import sys

import numpy as np
import pytest

import oracle
from omi_physics_amd import scenes
from omi_physics_amd.batch import batch_from_arrays

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
KEYS = ('position', 'orientation', 'linear_velocity', 'angular_velocity', 'sleep_timer')


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(x):
    return np.ascontiguousarray(x).view(np.uint8)


def run(arrays, precision, steps, island):
    b = batch_from_arrays([arrays], precision=precision, island_solve=island)
    b.step(DT, steps)
    b.sync()
    out = {'state': b.read(0), 'contacts': b.contacts(0), 'islands': b.islands(), 'levels': b.world_stats(1)['levels']}
    b.close()
    return out


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_island_solve_equals_world_solve(precision):
    """A 700-body pile, 150 steps: island by island and one block per world give
    identical states and contact impulses, byte for byte."""
    arrays = scenes.arrays(scenes.boxes_on_ground(700, seed=5, sphere_fraction=0.5))
    say('boxes on ground, 700 bodies, %s, 150 steps, islands vs world' % precision)
    a = run(arrays, precision, 150, True)
    w = run(arrays, precision, 150, False)
    assert a['islands']['size_class'] >= 0 and a['islands']['islands'] > 1, a['islands']
    assert w['islands']['size_class'] == -1
    for k in KEYS + ('awake',):
        assert np.array_equal(bits(a['state'][k]), bits(w['state'][k])), k
    for k in ('a', 'b', 'normal_impulse', 'tangent_impulse'):
        assert np.array_equal(bits(a['contacts'][k]), bits(w['contacts'][k])), k
    # the world's schedule is at least as deep as its deepest island's (the
    # islands here are small enough to be assigned pass by pass; the whole
    # world's last passes may be shifted copies of an earlier one)
    assert 0 < a['levels'][0] <= w['levels'][0]


def test_island_solve_fp64_matches_oracle():
    """Separate piles in one world (boxes dropped on a wide ground at low
    density, so many islands), 200 FP64 steps: states equal the oracle bit for
    bit at every check, while the step solves tens of islands."""
    arrays = scenes.arrays(scenes.boxes_on_ground(400, seed=11, sphere_fraction=0.3))
    say('boxes on ground, 400 bodies, fp64, 200 steps vs oracle')
    b = batch_from_arrays([arrays], precision='fp64')
    ow = oracle.OracleWorld(arrays)
    most = 0
    for _ in range(4):
        b.step(DT, 50)
        b.sync()
        ow.step(DT, 50)
        most = max(most, b.islands()['islands'])
        g, o = b.read(0), ow.state()
        for k in KEYS:
            assert np.array_equal(bits(g[k]), bits(o[k])), k
        assert np.array_equal(g['awake'], o['awake'])
    b.close()
    assert most > 10, most


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_clip16_pass_equals_register_clip(precision):
    """The narrowphase's 16-vertex clip pass (taken by a box pair whose clip
    outgrows the register polygon) routed every box/sphere pair through it on a
    300-body pile for 120 steps: states and contacts identical to the normal
    path byte for byte, and FP64 equal to the oracle."""
    arrays = scenes.arrays(scenes.boxes_on_ground(300, seed=8, spread=4.0, y_range=(0.6, 12.0),
                                                  sphere_fraction=0.5))
    say('clip16 pass for every pair, 300 bodies, %s, 120 steps' % precision)
    out = []
    for all16 in (False, True):
        b = batch_from_arrays([arrays], precision=precision, clip16_all=all16)
        b.step(DT, 120)
        b.sync()
        out.append((b.read(0), b.contacts(0)))
        b.close()
    (s0, c0), (s1, c1) = out
    assert len(c0['a']) > 100
    for k in KEYS + ('awake',):
        assert np.array_equal(bits(s0[k]), bits(s1[k])), k
    for k in ('a', 'b', 'point', 'normal', 'depth', 'normal_impulse', 'tangent_impulse'):
        assert np.array_equal(bits(c0[k]), bits(c1[k])), k
    if precision == 'fp64':
        ow = oracle.OracleWorld(arrays)
        ow.step(DT, 120)
        st = ow.state()
        for k in KEYS:
            assert np.array_equal(s1[k], st[k]), k
