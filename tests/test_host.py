"""Host-side logic and the C ABI surface (CPU only: no device calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from omi_physics_amd import PhysicsWorld, model, scenes, _ffi
from omi_physics_amd.world import inertia_diagonal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, 'include', 'omi_hip.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(omi_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_ffi.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_ffi.EXPORTS)


def test_binding_loads_without_gpu():
    L = _ffi.load()
    assert L.omi_abi_version() == 3
    from omi_physics_amd.batch import KERNELS
    assert tuple(L.omi_batch_kernel_names().decode().split(',')) == KERNELS


def test_fp64_solver_block_limit_rejected_before_device_use():
    # argument checks run before the library touches a device
    from omi_physics_amd.batch import Batch
    from omi_physics_amd.autotune import load_tuned
    with pytest.raises(RuntimeError, match='FP64 solver blocks'):
        Batch(4, 64, precision='fp64', solve_block=[64, 128, 512, 256])
    assert max(load_tuned('fp64')['solve_block']) <= 256


def test_mass_properties_follow_reference():
    w = PhysicsWorld()
    box = w.add_shape(model.Shape.box((1.0, 2.0, 3.0)))
    sph = w.add_shape(model.Shape.sphere(0.5))
    a = w.add_body(model.Motion(mass=2.0), collider=model.Collider(shape=box))
    b = w.add_body(model.Motion(mass=3.0), collider=model.Collider(shape=sph))
    c = w.add_body(model.Motion(type=model.STATIC), collider=model.Collider(shape=box))
    assert np.allclose(w.inv_mass, [0.5, 1 / 3, 0.0])
    k = 2.0 / 12.0
    assert np.allclose(w.inv_inertia[a], 1 / (k * np.array([4 + 9, 1 + 9, 1 + 4])))
    assert np.allclose(w.inv_inertia[b], 1 / (0.4 * 3.0 * 0.25))
    assert np.all(w.inv_inertia[c] == 0)
    assert np.allclose(inertia_diagonal(model.Shape.sphere(1.0), 1.0), 0.4)


def test_material_tables_match_reference_combine(golden):
    # the solver golden vectors carry the reference's per-contact restitution and friction
    rng = np.random.default_rng(0)
    w = PhysicsWorld()
    for f, e in ((0.5, 0.0), (0.8, 0.3), (0.2, 0.9)):
        w.add_material(model.Material(staticFriction=f, dynamicFriction=f, restitution=e))
    w.set_pair_friction(1, 2, 0.1, 0.05)
    box = w.add_shape(model.Shape.box((1, 1, 1)))
    for m in (-1, 0, 1, 2):
        w.add_body(model.Motion(), collider=model.Collider(shape=box, physicsMaterial=m))
    arr = w.export_arrays()
    s, nm = arr['mat_slot'], arr['n_mat']
    fr = arr['friction_table'].reshape(nm, nm); rs = arr['restitution_table'].reshape(nm, nm)
    # default (-1) with material 1 uses the default's average combine; 1-2 uses the override
    assert fr[s[0], s[2]] == 0.5 * (0.6 + 0.8)
    assert fr[s[2], s[3]] == 0.05 and fr[s[3], s[2]] == 0.05
    assert rs[s[2], s[3]] == 0.5 * (0.3 + 0.9)


def test_filter_table_is_collides_with():
    w = PhysicsWorld()
    f = w.add_filter(model.CollisionFilter(collisionSystems=('p',), notCollideWithSystems=('p',)))
    sh = w.add_shape(model.Shape.sphere(1.0))
    w.add_body(model.Motion(), collider=model.Collider(shape=sh, collisionFilter=f))
    w.add_body(model.Motion(), collider=model.Collider(shape=sh))
    arr = w.export_arrays()
    t = arr['filter_table'].reshape(arr['n_filt'], arr['n_filt'])
    s = arr['filt_slot']
    assert t[s[0], s[0]] == 0 and t[s[0], s[1]] == 1 and t[s[1], s[1]] == 1


def test_unsupported_shape_rejected():
    w = PhysicsWorld()
    with pytest.raises(ValueError):
        w.add_shape(model.Shape(type='capsule'))


def test_cpu_backend_names_raise():
    with pytest.raises(RuntimeError):
        PhysicsWorld(backend='numpy')


def test_scene_generators_are_seeded():
    a = scenes.boxes_on_ground(50, seed=4)
    b = scenes.boxes_on_ground(50, seed=4)
    assert a['bodies'] == b['bodies']
    sizes = scenes.world_sizes(4096)
    assert sizes.min() >= 8 and sizes.max() <= 2048 and 50 <= np.median(sizes) <= 80


def test_fast_arrays_equal_world_export():
    spec = scenes.boxes_on_ground(300, seed=9, sphere_fraction=0.5)
    a = scenes.arrays(spec)
    b = scenes.build(spec, PhysicsWorld, model).export_arrays()
    for key, val in b.items():
        if isinstance(val, np.ndarray):
            assert np.array_equal(np.asarray(a[key]).reshape(val.shape), val), key
        else:
            assert a[key] == val, key


def test_hull_vertex_order_matches_reference(golden):
    from omi_physics_amd.hull import convex_hull
    g = golden('hull.npz')
    co, ho = g['cloud_off'], g['hull_off']
    for i in range(len(co) - 1):
        v, tris = convex_hull(g['cloud_pts'][co[i]:co[i + 1]])
        assert np.array_equal(v, g['hull_pts'][ho[i]:ho[i + 1]]), i
        # closed, outward-oriented triangulation: Euler V - E + F = 2 and every point inside
        assert len(v) - 3 * len(tris) // 2 + len(tris) == 2
        c = v.mean(axis=0)
        for t in tris:
            n = np.cross(v[t[1]] - v[t[0]], v[t[2]] - v[t[0]])
            assert np.dot(n, c - v[t[0]]) < 0


def test_layout_of_a_host_column_world_equals_the_resident_world():
    """The C-ABI layout built from a reference-shaped host world (tests/seam_world.py)
    equals the one from PhysicsWorld and the vectorised scenes.arrays, for every
    column and table (trigger bodies and both kinds of shapes included)."""
    from omi_physics_amd.layout import world_arrays
    from tests import helpers
    from tests.seam_world import SeamWorld
    for spec in (helpers.trigger_scene(40, seed=2), helpers.pile(60, seed=4)):
        a = scenes.build(spec, PhysicsWorld, model).export_arrays()
        b = world_arrays(scenes.build(spec, SeamWorld, model), velocity_iterations=10)
        assert set(a) == set(b)
        for key, val in a.items():
            if isinstance(val, np.ndarray):
                assert np.array_equal(val, b[key]), key
            else:
                assert val == b[key], key


def test_resident_world_capacity_error_poisons_until_recovered():
    """A capacity error found when the mirror refreshes keeps the mirror's last good
    state, keeps the device marked newer, and fails every later read and step with the
    same message until recover_device(); then the next step reloads."""
    from omi_physics_amd.backend import DeviceWorld

    class FullBatch:
        def sync(self):
            raise RuntimeError('omi_hip error -3: batch capacity exceeded in a step since the last sync')

        def close(self):
            pass

    w = PhysicsWorld()
    w.add_body(model.Motion(), position=(1.0, 2.0, 3.0))
    dev = DeviceWorld('fp64', with_joints=True)
    dev.batch, dev.n, dev.key = FullBatch(), 1, ('layout',)
    w._hip_device = dev
    w._device_fresh = True
    for _ in range(2):
        with pytest.raises(RuntimeError, match='capacity'):
            w.position
        assert w._device_fresh and dev.error is not None
    with pytest.raises(RuntimeError, match='recover_device'):
        w.step(1.0 / 60.0)
    assert np.array_equal(w._raw('position')[0], [1.0, 2.0, 3.0])
    w.recover_device()
    assert dev.error is None and dev.batch is None and dev.grow == 2 and not w._device_fresh
    assert 'layout' in w._edited
