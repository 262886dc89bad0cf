"""omi_collide_convex_noalloc (device buffers, GJK/EPA workspaces in the library's
static scratch array) against omi_collide_convex on the same pairs, bit for bit:
two calls in a row, the second smaller, so the second reuses workspaces the
first left behind."""
import sys

import numpy as np
import pytest

from omi_physics_amd import _ffi
from omi_physics_amd._ffi import _d, _i32, check, ptr, c64

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hull_pairs(n, seed):
    """n pairs of cooked-ellipsoid hulls, boxes and spheres placed to overlap or nearly."""
    from omi_physics_amd.hull import cook_convex
    rng = np.random.default_rng(seed)
    hulls, off = [], [0]
    for _ in range(24):
        m = int(rng.integers(16, 33))
        d = rng.normal(size=(m, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        h = cook_convex(d * rng.uniform(0.3, 1.0, size=3))
        hulls.append(h)
        off.append(off[-1] + len(h))
    pts = np.concatenate(hulls)
    ka = rng.choice([1, 2, 3], size=n, p=[0.15, 0.15, 0.7]).astype(np.int32)
    kb = rng.choice([1, 2, 3], size=n, p=[0.15, 0.15, 0.7]).astype(np.int32)
    ia, ib = rng.integers(0, 24, size=n), rng.integers(0, 24, size=n)
    oa = np.where(ka == 3, np.array(off)[ia], 0).astype(np.int32)
    na = np.where(ka == 3, np.diff(off)[ia], 0).astype(np.int32)
    ob = np.where(kb == 3, np.array(off)[ib], 0).astype(np.int32)
    nb = np.where(kb == 3, np.diff(off)[ib], 0).astype(np.int32)

    def rot(k):
        q = rng.normal(size=(k, 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        x, y, z, w = q.T
        return np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                         2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                         2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=1).reshape(k, 3, 3)
    ca = rng.normal(scale=0.3, size=(n, 3))
    cb = ca + rng.normal(size=(n, 3)) * rng.uniform(0.3, 1.4, size=(n, 1))
    ha = np.where(ka[:, None] == 1, rng.uniform(0.2, 0.8, size=(n, 1)), rng.uniform(0.2, 0.8, size=(n, 3)))
    hb = np.where(kb[:, None] == 1, rng.uniform(0.2, 0.8, size=(n, 1)), rng.uniform(0.2, 0.8, size=(n, 3)))
    return dict(ka=ka, ca=ca, Ra=rot(n), ha=ha, oa=oa, na=na, kb=kb, cb=cb, Rb=rot(n), hb=hb, ob=ob, nb=nb, pts=pts)


def host_call(L, P):
    n = len(P['ka'])
    cnt = np.zeros(n, np.int32); nrm = np.zeros((n, 3)); pt = np.zeros((n, 3)); dep = np.zeros(n)
    keep = {k: (c64(v) if v.dtype == np.float64 else np.ascontiguousarray(v, np.int32)) for k, v in P.items()}
    a = lambda k: ptr(keep[k], _d if keep[k].dtype == np.float64 else _i32)
    check(L.omi_collide_convex(n, 64, a('ka'), a('ca'), a('Ra'), a('ha'), a('oa'), a('na'), a('kb'), a('cb'), a('Rb'),
                               a('hb'), a('ob'), a('nb'), a('pts'), len(P['pts']), ptr(cnt, _i32), ptr(nrm, _d),
                               ptr(pt, _d), ptr(dep, _d)))
    return cnt, nrm, pt, dep


class DeviceBuffers:
    """Device copies of host arrays through the HIP runtime libomi_hip.so itself
    links (torch bundles another runtime, whose pointers this library must not
    be handed).  The caller allocates and frees; the entry point does neither."""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.hip = ctypes.CDLL('/opt/rocm/lib/libamdhip64.so.7')
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.ct.c_void_p()
        assert self.hip.hipMalloc(self.ct.byref(p), max(a.nbytes, 8)) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0          # host to device
        self.live.append(p)
        return p.value

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0      # device to host
        return out

    def free(self):
        for p in self.live:
            self.hip.hipFree(p)
        self.live = []


def device_call(L, P):
    n = len(P['ka'])
    D = DeviceBuffers()
    try:
        T = {k: D.put(v) for k, v in P.items()}
        outs = [np.full(n, 7, np.int32), np.full((n, 3), 7.0), np.full((n, 3), 7.0), np.full(n, 7.0)]
        dp = [D.put(o) for o in outs]
        check(L.omi_collide_convex_noalloc(n, T['ka'], T['ca'], T['Ra'], T['ha'], T['oa'], T['na'], T['kb'], T['cb'],
                                           T['Rb'], T['hb'], T['ob'], T['nb'], T['pts'], len(P['pts']), dp[0], dp[1],
                                           dp[2], dp[3], None))
        assert D.hip.hipDeviceSynchronize() == 0
        return [D.get(p, o) for p, o in zip(dp, outs)]
    finally:
        D.free()


def test_noalloc_convex_equals_allocating_entry_point():
    L = _ffi.load()
    if L.omi_device_count() < 1:
        pytest.fail('no HIP device visible')
    for n, seed in ((3000, 1), (700, 2)):           # 3000 pairs reuse the 512 workspaces; then fewer pairs
        P = hull_pairs(n, seed)
        say('noalloc convex: %d pairs (seed %d)' % (n, seed))
        ref = host_call(L, P)
        got = device_call(L, P)
        hits = int((ref[0] == 1).sum())
        say('  contacts %d of %d' % (hits, n))
        assert hits > n // 10
        for x, y in zip(got, ref):
            assert np.array_equal(x, y)


def test_noalloc_convex_flags_bad_pairs_and_rejects_bad_args():
    """A pair with an unknown kind or a hull range beyond the points gets count -1 and
    zero outputs (no hull read); a negative pair count or a NULL buffer is an argument
    error before any launch."""
    L = _ffi.load()
    P = hull_pairs(64, 3)
    P['ka'][5] = 9
    P['oa'][7] = len(P['pts']) - 2; P['ka'][7] = 3; P['na'][7] = 16
    P['kb'][9] = 3; P['ob'][9] = -4; P['nb'][9] = 16
    cnt, nrm, pt, dep = device_call(L, P)
    for i in (5, 7, 9):
        assert cnt[i] == -1 and not nrm[i].any() and not pt[i].any() and dep[i] == 0.0
    ok = np.ones(64, bool); ok[[5, 7, 9]] = False
    Q = {k: (v if k == 'pts' else v[ok]) for k, v in P.items()}
    ref = host_call(L, Q)
    assert np.array_equal(cnt[ok], ref[0]) and np.array_equal(dep[ok], ref[3])
    D = DeviceBuffers()
    try:
        x = D.put(np.zeros(8))
        assert L.omi_collide_convex_noalloc(-1, *([x] * 13), 8, *([x] * 4), None) == _ffi.OMI_ERR_ARG
        assert L.omi_collide_convex_noalloc(4, *([x] * 12), None, 8, *([x] * 4), None) == _ffi.OMI_ERR_ARG
    finally:
        D.free()
