"""Device-resident batches of independent worlds (the batched form of the step).

One :class:`Batch` owns one ``omi_batch``: every loaded world's state stays in
HBM, and :meth:`Batch.step` enqueues whole steps for all of them at once —
per-body kernels over every body, one workgroup per world for the
sweep-and-prune and the solver — captured in a HIP graph after the first step.
A single world is simply a batch of one (the plain form).  A
:class:`BatchGroup` splits worlds of different size classes over member batches
whose steps run side by side.
"""
import ctypes
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _ffi
from ._ffi import _d, _i32, _i64, _u8, check, ptr
from .autotune import size_class

KERNELS = ('begin', 'pre', 'sap', 'narrow', 'compact', 'prepare', 'solve_top', 'solve_rest', 'cache',
           'joints', 'post', 'sleep', 'toggle')


@dataclass
class Contact:
    """One contact point (reference ``collide.py:37-53``): normal a -> b, depth > 0."""
    a: int
    b: int
    point: np.ndarray
    normal: np.ndarray
    depth: float
    normal_impulse: float = 0.0
    tangent_impulse: np.ndarray = field(default_factory=lambda: np.zeros(2))


def capacities_for(n_bodies: int, n_worlds: int = 1, pairs_per_body: float = 12.0,
                   contacts_per_body: float = 24.0) -> Dict[str, int]:
    """Pair/contact capacities that cover piles of boxes and spheres with margin (within the
    library's limits of 2^30 pairs and 2^24 contacts per step)."""
    return dict(max_pairs=min(int(pairs_per_body * n_bodies) + 4096 * max(1, n_worlds // 64 + 1), 1 << 30),
                max_contacts=min(int(contacts_per_body * n_bodies) + 8192 * max(1, n_worlds // 64 + 1), 1 << 24))


class Batch:
    """Many worlds stepped together on one GPU."""

    def __init__(self, max_worlds: int, max_bodies: int, precision: str = 'fp64',
                 max_pairs: Optional[int] = None, max_contacts: Optional[int] = None,
                 max_hull_pts: int = 1, max_joints: int = 1, record_broadphase: bool = False,
                 solver_block: int = 0, use_graph: bool = True,
                 sap_block: Optional[Sequence[int]] = None,
                 solve_block: Optional[Sequence[int]] = None, island_solve: bool = True,
                 clip16_all: bool = False) -> None:
        self.lib = _ffi.load()
        caps = capacities_for(max_bodies, max_worlds)
        cfg = _ffi.BatchConfig()
        cfg.max_worlds = int(max_worlds)
        cfg.precision = _ffi.OMI_FP64 if precision in ('fp64', 64) else _ffi.OMI_FP32
        cfg.max_bodies = int(max_bodies)
        cfg.max_pairs = int(max_pairs or caps['max_pairs'])
        cfg.max_contacts = int(max_contacts or caps['max_contacts'])
        cfg.max_hull_pts = int(max(1, max_hull_pts))
        cfg.max_joints = int(max(1, max_joints))
        cfg.record_broadphase = int(bool(record_broadphase))
        cfg.solver_block = int(solver_block)
        cfg.island_solve = 0 if island_solve else -1
        cfg.clip16_all = int(bool(clip16_all))
        from .autotune import load_tuned
        tuned = load_tuned(precision)
        for c in range(4):
            cfg.sap_block[c] = int((sap_block or tuned['sap_block'])[c])
            cfg.solve_block[c] = int((solve_block or tuned['solve_block'])[c])
        self.config = cfg
        self.precision = 'fp64' if cfg.precision == 64 else 'fp32'
        self.max_worlds = int(max_worlds)
        h = ctypes.c_void_p()
        check(self.lib.omi_batch_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.sizes: Dict[int, int] = {}
        check(self.lib.omi_batch_use_graph(self._h, int(bool(use_graph))))

    def close(self) -> None:
        if getattr(self, '_h', None):
            self.lib.omi_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- worlds -------------------------------------------------------------
    def load(self, slot: int, arrays: dict) -> None:
        """Place world ``arrays`` (``export_world_arrays`` output) in ``slot``."""
        desc, keep = _ffi.world_desc(arrays)
        check(self.lib.omi_batch_load_world(self._h, int(slot), ctypes.byref(desc)))
        self.sizes[int(slot)] = int(arrays['n'])

    def remove(self, slot: int) -> None:
        check(self.lib.omi_batch_remove_world(self._h, int(slot)))
        self.sizes.pop(int(slot), None)

    def set_state(self, slot: int, position=None, orientation=None, linear_velocity=None,
                  angular_velocity=None, awake=None, sleep_timer=None) -> None:
        keep = [None if x is None else _ffi.c64(x) for x in (position, orientation, linear_velocity,
                                                              angular_velocity)]
        aw = None if awake is None else np.ascontiguousarray(awake, dtype=np.uint8)
        tm = None if sleep_timer is None else _ffi.c64(sleep_timer)
        check(self.lib.omi_batch_set_state(self._h, int(slot), ptr(keep[0], _d), ptr(keep[1], _d),
                                           ptr(keep[2], _d), ptr(keep[3], _d), ptr(aw, _u8), ptr(tm, _d)))

    def set_aabbs(self, slot: int, aabb_min=None, aabb_max=None) -> None:
        lo = None if aabb_min is None else _ffi.c64(aabb_min)
        hi = None if aabb_max is None else _ffi.c64(aabb_max)
        check(self.lib.omi_batch_set_aabbs(self._h, int(slot), ptr(lo, _d), ptr(hi, _d)))

    def set_params(self, slot: int, inv_mass=None, inv_inertia=None, gravity_factor=None, linear_damping=None,
                   angular_damping=None, quadratic_drag=None, gravity=None, default_damping=None,
                   sleep_enabled: Optional[bool] = None, velocity_iterations: Optional[int] = None) -> None:
        """Overwrite per-body parameters and world settings of a loaded world (None keeps a value)."""
        arrs = [None if x is None else _ffi.c64(x) for x in (inv_mass, inv_inertia, gravity_factor, linear_damping,
                                                             angular_damping, quadratic_drag, gravity,
                                                             default_damping)]
        check(self.lib.omi_batch_set_params(
            self._h, int(slot), *[ptr(a, _d) for a in arrs],
            -1 if sleep_enabled is None else int(bool(sleep_enabled)),
            -1 if velocity_iterations is None else int(velocity_iterations)))

    def read_cache(self, slot: int) -> Dict[str, np.ndarray]:
        """One world's warm-start cache (world-local pairs, points, impulses)."""
        cnt = np.zeros(1, np.int64)
        check(self.lib.omi_batch_read_cache(self._h, int(slot), 0, ptr(cnt, _i64), None, None, None, None, None))
        K = int(cnt[0])
        out = dict(a=np.zeros(K, np.int32), b=np.zeros(K, np.int32), point=np.zeros((K, 3)),
                   normal_impulse=np.zeros(K), tangent_impulse=np.zeros((K, 2)))
        if K:
            check(self.lib.omi_batch_read_cache(self._h, int(slot), K, ptr(cnt, _i64), ptr(out['a'], _i32),
                                                ptr(out['b'], _i32), ptr(out['point'], _d),
                                                ptr(out['normal_impulse'], _d), ptr(out['tangent_impulse'], _d)))
        return out

    def write_cache(self, slot: int, cache: Dict[str, np.ndarray]) -> None:
        """Replace one world's warm-start cache (the layout :meth:`read_cache` returns)."""
        a = np.ascontiguousarray(cache['a'], np.int32)
        b = np.ascontiguousarray(cache['b'], np.int32)
        pt, ni, ti = (_ffi.c64(cache[k]) for k in ('point', 'normal_impulse', 'tangent_impulse'))
        check(self.lib.omi_batch_write_cache(self._h, int(slot), len(a), ptr(a, _i32), ptr(b, _i32), ptr(pt, _d),
                                             ptr(ni, _d), ptr(ti, _d)))

    def read_joint_anchors(self, slot: int, n_joints: int):
        """``(local_a, local_b, fixed)`` of one world's joints as the device holds them."""
        la, lb, fx = np.zeros((n_joints, 3)), np.zeros((n_joints, 3)), np.zeros(n_joints, np.uint8)
        if n_joints:
            check(self.lib.omi_batch_read_joint_anchors(self._h, int(slot), ptr(la, _d), ptr(lb, _d), ptr(fx, _u8)))
        return la, lb, fx

    # -- stepping -----------------------------------------------------------
    def collide(self, dt: float) -> None:
        """Enqueue the collision stage alone for every loaded world (asynchronous)."""
        check(self.lib.omi_batch_collide(self._h, float(dt)))

    def step(self, dt: float, steps: int = 1) -> None:
        """Enqueue ``steps`` steps of ``dt`` for every loaded world (asynchronous)."""
        check(self.lib.omi_batch_step(self._h, float(dt), int(steps)))

    def prepare(self, dt: float) -> None:
        """Capture the step graph for ``dt`` ahead of the first step."""
        check(self.lib.omi_batch_prepare(self._h, float(dt)))

    def refit(self, margin: float = 0.05) -> None:
        check(self.lib.omi_batch_refit(self._h, float(margin)))

    def sync(self) -> None:
        """Wait for queued steps; raises if a step ran out of pair/contact capacity."""
        check(self.lib.omi_batch_sync(self._h))

    def use_graph(self, on: bool) -> None:
        check(self.lib.omi_batch_use_graph(self._h, int(bool(on))))

    def set_blocks(self, sap_block: Sequence[int], solve_block: Sequence[int]) -> None:
        """Threads per world of the broadphase sort / solver per size class (0 keeps a value)."""
        sb = np.ascontiguousarray(sap_block, dtype=np.int32)
        vb = np.ascontiguousarray(solve_block, dtype=np.int32)
        check(self.lib.omi_batch_set_blocks(self._h, ptr(sb, _i32), ptr(vb, _i32)))

    def set_timing(self, on: bool) -> None:
        check(self.lib.omi_batch_set_timing(self._h, int(bool(on))))

    def kernel_times(self) -> Dict[str, float]:
        """Per-stage device milliseconds of the last timed step."""
        ms = np.zeros(len(KERNELS))
        check(self.lib.omi_batch_kernel_times(self._h, len(KERNELS), ptr(ms, _d)))
        return dict(zip(KERNELS, ms.tolist()))

    # -- readback -----------------------------------------------------------
    def read(self, slot: int) -> Dict[str, np.ndarray]:
        """Copies of one world's mutable state."""
        n = self.sizes[int(slot)]
        out = dict(position=np.empty((n, 3)), orientation=np.empty((n, 4)),
                   linear_velocity=np.empty((n, 3)), angular_velocity=np.empty((n, 3)),
                   awake=np.empty(n, np.uint8), sleep_timer=np.empty(n),
                   aabb_min=np.empty((n, 3)), aabb_max=np.empty((n, 3)),
                   prev_position=np.empty((n, 3)), prev_orientation=np.empty((n, 4)))
        check(self.lib.omi_batch_read_world(
            self._h, int(slot), ptr(out['position'], _d), ptr(out['orientation'], _d),
            ptr(out['linear_velocity'], _d), ptr(out['angular_velocity'], _d), ptr(out['awake'], _u8),
            ptr(out['sleep_timer'], _d), ptr(out['aabb_min'], _d), ptr(out['aabb_max'], _d),
            ptr(out['prev_position'], _d), ptr(out['prev_orientation'], _d)))
        return out

    def mark(self) -> int:
        """A marker after the queued work (asynchronous); see :meth:`mark_wait`."""
        t = np.zeros(1, np.int64)
        check(self.lib.omi_batch_mark(self._h, ptr(t, _i64)))
        return int(t[0])

    def mark_wait(self, ticket: int) -> None:
        """Wait for the work queued before the marker; raises on a capacity overflow
        of those steps (the flags stay set until :meth:`sync`)."""
        check(self.lib.omi_batch_mark_wait(self._h, int(ticket)))

    def fetch(self, slot: int):
        """Queue a copy of one world's state after the queued steps; returns a ticket
        for :meth:`fetched` (the device is not waited for)."""
        t = np.zeros(1, np.int64)
        check(self.lib.omi_batch_fetch_world(self._h, int(slot), ptr(t, _i64)))
        return int(t[0]), self.sizes[int(slot)]

    def fetch_ready(self, ticket) -> bool:
        r = int(self.lib.omi_batch_fetch_ready(self._h, int(ticket[0])))
        if r < 0:
            check(r)
        return r == 1

    def fetched(self, ticket) -> Dict[str, np.ndarray]:
        """The state a :meth:`fetch` copied (waits for it; the layout of :meth:`read`)."""
        t, n = ticket
        out = dict(position=np.empty((n, 3)), orientation=np.empty((n, 4)),
                   linear_velocity=np.empty((n, 3)), angular_velocity=np.empty((n, 3)),
                   awake=np.empty(n, np.uint8), sleep_timer=np.empty(n),
                   aabb_min=np.empty((n, 3)), aabb_max=np.empty((n, 3)),
                   prev_position=np.empty((n, 3)), prev_orientation=np.empty((n, 4)))
        check(self.lib.omi_batch_fetch_wait(
            self._h, int(t), ptr(out['position'], _d), ptr(out['orientation'], _d),
            ptr(out['linear_velocity'], _d), ptr(out['angular_velocity'], _d), ptr(out['awake'], _u8),
            ptr(out['sleep_timer'], _d), ptr(out['aabb_min'], _d), ptr(out['aabb_max'], _d),
            ptr(out['prev_position'], _d), ptr(out['prev_orientation'], _d)))
        return out

    def snapshot(self, slots: Optional[Sequence[int]] = None) -> List[tuple]:
        """Pose snapshots ``(position (n,3), axis_angle (n,4), awake (n,), dynamic (n,))``
        of ``slots`` (default: every loaded world), in the reference's snapshot format
        (``threaded.py:16-21``), packed on the device and downloaded at once."""
        slots = sorted(self.sizes) if slots is None else [int(s) for s in slots]
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        off = np.zeros(len(slots) + 1, np.int64)
        check(self.lib.omi_batch_snapshot(self._h, len(slots), ptr(sl, _i32), ptr(off, _i64), None, None,
                                          None, None))
        n = int(off[-1])
        pos, aa = np.empty((n, 3)), np.empty((n, 4))
        aw, dy = np.empty(n, np.uint8), np.empty(n, np.uint8)
        check(self.lib.omi_batch_snapshot(self._h, len(slots), ptr(sl, _i32), ptr(off, _i64), ptr(pos, _d),
                                          ptr(aa, _d), ptr(aw, _u8), ptr(dy, _u8)))
        return [(pos[a:b], aa[a:b], aw[a:b].astype(bool), dy[a:b].astype(bool))
                for a, b in zip(off[:-1], off[1:])]

    def contacts(self, slot: int) -> Dict[str, np.ndarray]:
        """One world's last-step contacts in solver order (world-local indices)."""
        cnt = np.zeros(1, np.int64)
        check(self.lib.omi_batch_read_contacts(self._h, int(slot), 0, ptr(cnt, _i64), None, None, None,
                                               None, None, None, None))
        K = int(cnt[0])
        out = dict(a=np.empty(K, np.int32), b=np.empty(K, np.int32), point=np.empty((K, 3)),
                   normal=np.empty((K, 3)), depth=np.empty(K), normal_impulse=np.empty(K),
                   tangent_impulse=np.empty((K, 2)))
        if K:
            check(self.lib.omi_batch_read_contacts(
                self._h, int(slot), K, ptr(cnt, _i64), ptr(out['a'], _i32), ptr(out['b'], _i32),
                ptr(out['point'], _d), ptr(out['normal'], _d), ptr(out['depth'], _d),
                ptr(out['normal_impulse'], _d), ptr(out['tangent_impulse'], _d)))
        return out

    def trigger_overlaps(self, slot: int) -> np.ndarray:
        """One world's last-step trigger overlaps ``(k, 2)``: (trigger body, other body),
        world-local, sorted."""
        cnt = np.zeros(1, np.int64)
        check(self.lib.omi_batch_read_trigger_overlaps(self._h, int(slot), 0, ptr(cnt, _i64), None, None))
        k = int(cnt[0])
        t, o = np.empty(k, np.int32), np.empty(k, np.int32)
        if k:
            check(self.lib.omi_batch_read_trigger_overlaps(self._h, int(slot), k, ptr(cnt, _i64), ptr(t, _i32),
                                                           ptr(o, _i32)))
        return np.stack([t, o], axis=1) if k else np.zeros((0, 2), np.int32)

    def pairs(self, slot: int, solid: bool = True) -> np.ndarray:
        """One world's last-step pairs ``(P, 2)`` (solid = narrowphase input)."""
        cnt = np.zeros(1, np.int64)
        check(self.lib.omi_batch_read_pairs(self._h, int(slot), int(solid), 0, ptr(cnt, _i64), None))
        P = int(cnt[0])
        out = np.empty((P, 2), np.int32)
        if P:
            check(self.lib.omi_batch_read_pairs(self._h, int(slot), int(solid), P, ptr(cnt, _i64),
                                                ptr(out, _i32)))
        return out

    def counters(self) -> Dict[str, int]:
        """Totals of the last step over all worlds."""
        o = np.zeros(8, np.int64)
        check(self.lib.omi_batch_counters(self._h, ptr(o, _i64)))
        keys = ('bodies', 'x_candidates', 'broadphase_pairs', 'solid_pairs', 'manifolds', 'contacts',
                'max_levels', 'overflow')
        return {k: int(v) for k, v in zip(keys, o)}

    def world_stats(self, n: Optional[int] = None) -> Dict[str, np.ndarray]:
        """Per-world figures of the last step for slots [0, n): bodies, contacts, manifolds,
        solver levels and size class (-1 for a free slot)."""
        n = self.max_worlds if n is None else int(n)
        out = {k: np.zeros(n, np.int32) for k in ('bodies', 'contacts', 'manifolds', 'levels', 'size_class')}
        check(self.lib.omi_batch_world_stats(self._h, n, *(ptr(out[k], _i32) for k in
                                                            ('bodies', 'contacts', 'manifolds', 'levels',
                                                             'size_class'))))
        return out

    def islands(self) -> Dict[str, int]:
        """The size class solved island by island (-1: none) and its islands in the last step."""
        c = np.zeros(1, np.int32)
        k = np.zeros(1, np.int64)
        check(self.lib.omi_batch_islands(self._h, ptr(c, _i32), ptr(k, _i64)))
        return {'size_class': int(c[0]), 'islands': int(k[0])}

    def device_bytes(self) -> int:
        return int(self.lib.omi_batch_device_bytes(self._h))


def _single_batch(worlds: Sequence[dict], precision: str, **kw) -> Batch:
    if 'sap_block' not in kw and 'solve_block' not in kw:
        from .autotune import blocks_for
        kw.update(blocks_for([int(w['n']) for w in worlds], precision))
    total = int(sum(int(w['n']) for w in worlds))
    hull = int(sum(len(np.asarray(w['hull_pts']).reshape(-1, 3)) for w in worlds))
    joints = int(sum(int(w.get('n_joints', 0)) for w in worlds))
    b = Batch(max_worlds=max(len(worlds), 1), max_bodies=max(total, 1), precision=precision,
              max_hull_pts=max(hull, 1), max_joints=max(joints, 1), **kw)
    for i, w in enumerate(worlds):
        b.load(i, w)
    return b


class BatchGroup:
    """Worlds split over member batches (lanes) by size class and stepped together.

    Each lane's step is an independent branch of one captured step (``omi_group_*``),
    so the few largest worlds solve while the lanes of smaller worlds run their
    chip-wide stages: a world's step no longer waits at stage boundaries shared with
    every other world.  Worlds are independent, so every world's results equal those of
    a single batch.  World ``i`` keeps index ``i`` in every method.
    """

    def __init__(self, worlds: Sequence[dict], lanes: Sequence[Sequence[int]], precision: str = 'fp64',
                 **kw) -> None:
        self.lib = _ffi.load()
        self.precision = 'fp64' if precision in ('fp64', 64) else 'fp32'
        self.lanes: List[Batch] = []
        self.lane_worlds: List[List[int]] = [list(map(int, l)) for l in lanes if len(l)]
        self.where: Dict[int, tuple] = {}
        self._h = None
        try:
            for li, idx in enumerate(self.lane_worlds):
                self.lanes.append(_single_batch([worlds[i] for i in idx], precision, **kw))
                for slot, i in enumerate(idx):
                    self.where[i] = (li, slot)
            arr = (ctypes.c_void_p * len(self.lanes))(*[b._h.value for b in self.lanes])
            h = ctypes.c_void_p()
            check(self.lib.omi_group_create(arr, len(self.lanes), ctypes.byref(h)))
            self._h = h
        except Exception:
            self.close()
            raise
        self.max_worlds = len(worlds)
        self.sizes = {i: int(worlds[i]['n']) for i in range(len(worlds))}

    def close(self) -> None:
        if getattr(self, '_h', None):
            self.lib.omi_group_destroy(self._h)
            self._h = None
        for b in getattr(self, 'lanes', []):
            b.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _at(self, i: int):
        li, slot = self.where[int(i)]
        return self.lanes[li], slot

    # -- stepping -----------------------------------------------------------
    def step(self, dt: float, steps: int = 1) -> None:
        check(self.lib.omi_group_step(self._h, float(dt), int(steps)))

    def prepare(self, dt: float) -> None:
        check(self.lib.omi_group_prepare(self._h, float(dt)))

    def sync(self) -> None:
        check(self.lib.omi_group_sync(self._h))

    def refit(self, margin: float = 0.05) -> None:
        for b in self.lanes:
            b.refit(margin)

    def use_graph(self, on: bool) -> None:
        for b in self.lanes:
            b.use_graph(on)

    def set_blocks(self, sap_block: Sequence[int], solve_block: Sequence[int]) -> None:
        for b in self.lanes:
            b.set_blocks(sap_block, solve_block)

    def set_timing(self, on: bool) -> None:
        for b in self.lanes:
            b.set_timing(on)

    def lane_kernel_times(self) -> List[Dict[str, float]]:
        """Per-stage device milliseconds of the last timed step, one dictionary per lane."""
        return [b.kernel_times() for b in self.lanes]

    # -- worlds -------------------------------------------------------------
    def set_state(self, i: int, **state) -> None:
        b, slot = self._at(i)
        b.set_state(slot, **state)

    def read(self, i: int) -> Dict[str, np.ndarray]:
        b, slot = self._at(i)
        return b.read(slot)

    def snapshot(self, worlds: Optional[Sequence[int]] = None) -> List[tuple]:
        """As :meth:`Batch.snapshot`, indexed by world (one packed download per lane)."""
        worlds = list(range(self.max_worlds)) if worlds is None else [int(i) for i in worlds]
        out: List[Optional[tuple]] = [None] * len(worlds)
        for li, b in enumerate(self.lanes):
            pick = [(k, self.where[i][1]) for k, i in enumerate(worlds) if self.where[i][0] == li]
            if pick:
                for (k, _), snap in zip(pick, b.snapshot([s for _, s in pick])):
                    out[k] = snap
        return out

    def contacts(self, i: int) -> Dict[str, np.ndarray]:
        b, slot = self._at(i)
        return b.contacts(slot)

    def pairs(self, i: int, solid: bool = True) -> np.ndarray:
        b, slot = self._at(i)
        return b.pairs(slot, solid)

    def trigger_overlaps(self, i: int) -> np.ndarray:
        b, slot = self._at(i)
        return b.trigger_overlaps(slot)

    def counters(self) -> Dict[str, int]:
        """Totals of the last step over all lanes (``max_levels`` the largest)."""
        out: Dict[str, int] = {}
        for b in self.lanes:
            for k, v in b.counters().items():
                out[k] = max(out.get(k, 0), v) if k == 'max_levels' else (out.get(k, 0) | v if k == 'overflow'
                                                                           else out.get(k, 0) + v)
        return out

    def lane_counters(self) -> List[Dict[str, int]]:
        return [b.counters() for b in self.lanes]

    def world_stats(self, n: Optional[int] = None) -> Dict[str, np.ndarray]:
        """As :meth:`Batch.world_stats`, indexed by world."""
        n = self.max_worlds if n is None else int(n)
        keys = ('bodies', 'contacts', 'manifolds', 'levels', 'size_class')
        out = {k: np.zeros(n, np.int32) for k in keys}
        out['size_class'][:] = -1
        for b, idx in zip(self.lanes, self.lane_worlds):
            st = b.world_stats(len(idx))
            sel = [(slot, i) for slot, i in enumerate(idx) if i < n]
            if sel:
                s_, i_ = np.array(sel).T
                for k in keys:
                    out[k][i_] = st[k][s_]
        return out

    def device_bytes(self) -> int:
        return int(sum(b.device_bytes() for b in self.lanes))


def batch_from_arrays(worlds: Sequence[dict], precision: str = 'fp64', lanes: bool = False, **kw):
    """A batch holding ``worlds`` as worlds 0..len-1, sized for them.

    With ``lanes`` the worlds are split by size class into the lanes of a
    :class:`BatchGroup` (when they span more than one class); otherwise one
    :class:`Batch` holds them all.
    """
    classes = [size_class(int(w['n'])) for w in worlds]
    if lanes and len(set(classes)) < 2:
        lanes = False
    if not lanes:
        return _single_batch(worlds, precision, **kw)
    groups = [[i for i, c in enumerate(classes) if c == k] for k in (3, 2, 1, 0)]
    return BatchGroup(worlds, [g for g in groups if g], precision=precision, **kw)


def contact_list(c: Dict[str, np.ndarray]) -> List[Contact]:
    """Contact objects from a ``Batch.contacts`` dictionary."""
    return [Contact(int(c['a'][k]), int(c['b'][k]), c['point'][k].copy(), c['normal'][k].copy(),
                    float(c['depth'][k]), float(c['normal_impulse'][k]), c['tangent_impulse'][k].copy())
            for k in range(len(c['a']))]
