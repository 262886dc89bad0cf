"""Continuous batching of world rollouts on one GPU.

Requests (a world plus a number of steps) arrive over time.  The server keeps
one device batch whose body/pair/contact capacity is the memory budget; each
tick it admits waiting requests into free world slots while their bodies fit,
advances every resident world by one step in one graph-launched step, and
retires finished worlds (reading their final state back and freeing their
slots for the next admissions).  A world's state never leaves HBM between its
admission and its retirement, and worlds admitted mid-flight share the step
launches of the worlds already running.

Per-world results do not depend on the mix: every kernel works on one world's
bodies, pairs and contacts only, so a world stepped in the server matches the
same world stepped alone bit for bit (``tests/test_gpu_serve.py``).
"""
import collections
import time
from dataclasses import dataclass, field
from typing import Any, Deque, Dict, List, Optional

import numpy as np

from . import _ffi
from .batch import Batch, capacities_for


@dataclass
class Request:
    """One rollout: ``arrays`` (``export_world_arrays`` layout), ``steps`` at ``dt``."""
    id: int
    arrays: Dict[str, Any]
    steps: int
    arrival: float = 0.0
    # filled by the server
    slot: int = -1
    done_steps: int = 0
    admitted_at: float = -1.0
    finished_at: float = -1.0
    result: Optional[Dict[str, np.ndarray]] = None

    @property
    def bodies(self) -> int:
        return int(self.arrays['n'])


@dataclass
class ServeStats:
    step_latency_s: List[float] = field(default_factory=list)
    active_bodies: List[int] = field(default_factory=list)
    busy_s: float = 0.0
    wall_s: float = 0.0

    def summary(self) -> Dict[str, float]:
        lat = np.array(self.step_latency_s) if self.step_latency_s else np.zeros(1)
        bs = float(np.sum(self.active_bodies))
        return {'ticks': len(self.step_latency_s),
                'body_steps': bs,
                'body_steps_per_s': bs / self.busy_s if self.busy_s > 0 else 0.0,
                'p50_step_ms': 1e3 * float(np.percentile(lat, 50)),
                'p99_step_ms': 1e3 * float(np.percentile(lat, 99)),
                'mean_active_bodies': float(np.mean(self.active_bodies)) if self.active_bodies else 0.0,
                'busy_s': self.busy_s, 'wall_s': self.wall_s}


def device_memory() -> Dict[str, int]:
    """Free and total HBM bytes of the current device."""
    L = _ffi.load()
    free = np.zeros(1, np.int64)
    total = np.zeros(1, np.int64)
    _ffi.check(L.omi_device_memory(_ffi.ptr(free, _ffi._i64), _ffi.ptr(total, _ffi._i64)))
    return {'free': int(free[0]), 'total': int(total[0])}


class RolloutServer:
    """Admits worlds into one resident batch as slots and memory free up."""

    def __init__(self, max_worlds: int = 512, max_bodies: Optional[int] = None,
                 precision: str = 'fp32', dt: float = 1.0 / 60.0,
                 memory_fraction: float = 0.25, max_hull_pts: int = 1) -> None:
        if max_bodies is None:
            # size the arena from the HBM budget: ~2.5 KB of device memory per body of capacity
            budget = int(device_memory()['free'] * memory_fraction)
            max_bodies = max(4096, budget // 2560)
        caps = capacities_for(max_bodies, max_worlds)
        self.batch = Batch(max_worlds, max_bodies, precision=precision, max_hull_pts=max_hull_pts, **caps)
        self.max_worlds = max_worlds
        self.max_bodies = max_bodies
        self.dt = dt
        self.free_slots: Deque[int] = collections.deque(range(max_worlds))
        self.resident: Dict[int, Request] = {}
        self.waiting: Deque[Request] = collections.deque()
        self.finished: List[Request] = []
        self.bodies_in_use = 0
        self.stats = ServeStats()

    def close(self) -> None:
        self.batch.close()

    def submit(self, req: Request) -> None:
        if req.bodies > self.max_bodies:
            raise ValueError('world of %d bodies exceeds the server capacity %d' % (req.bodies, self.max_bodies))
        self.waiting.append(req)

    def _admit(self, now: float) -> int:
        admitted = 0
        while self.waiting and self.free_slots and self.waiting[0].arrival <= now:
            req = self.waiting[0]
            if self.bodies_in_use + req.bodies > self.max_bodies:
                break                      # FIFO: wait for memory rather than overtaking
            slot = self.free_slots[0]
            try:
                self.batch.load(slot, req.arrays)
            except RuntimeError as e:
                if 'capacity' in str(e):
                    break                  # fragmented arena: wait for retirements
                raise
            self.waiting.popleft()
            self.free_slots.popleft()
            req.slot = slot
            req.admitted_at = now
            self.resident[slot] = req
            self.bodies_in_use += req.bodies
            admitted += 1
        return admitted

    def _retire(self, now: float) -> None:
        for slot in [s for s, r in self.resident.items() if r.done_steps >= r.steps]:
            req = self.resident.pop(slot)
            req.result = self.batch.read(slot)
            req.finished_at = now
            self.batch.remove(slot)
            self.free_slots.append(slot)
            self.bodies_in_use -= req.bodies
            self.finished.append(req)

    def tick(self, now: float) -> bool:
        """Admit, step every resident world once, retire; False when idle."""
        self._admit(now)
        if not self.resident:
            return False
        t = time.perf_counter()
        self.batch.step(self.dt, 1)
        self.batch.sync()
        dt = time.perf_counter() - t
        self.stats.step_latency_s.append(dt)
        self.stats.active_bodies.append(self.bodies_in_use)
        self.stats.busy_s += dt
        for r in self.resident.values():
            r.done_steps += 1
        self._retire(now)
        return True

    def run(self, requests: List[Request], realtime: bool = True) -> ServeStats:
        """Serve ``requests`` (sorted by arrival); with ``realtime`` the clock is
        wall time since start, idle gaps are skipped ahead rather than slept."""
        for r in sorted(requests, key=lambda r: r.arrival):
            self.submit(r)
        t0 = time.perf_counter()
        skew = 0.0
        self.batch.prepare(self.dt)
        while self.waiting or self.resident:
            now = (time.perf_counter() - t0) + skew if realtime else float('inf')
            if not self.tick(now) and self.waiting:
                skew += max(0.0, self.waiting[0].arrival - now)   # idle: jump to the next arrival
        self.stats.wall_s = time.perf_counter() - t0
        return self.stats


def run_plain(requests: List[Request], precision: str = 'fp32', dt: float = 1.0 / 60.0) -> Dict[str, float]:
    """The plain form: each request alone on the GPU, one after another."""
    from .batch import batch_from_arrays
    lat: List[float] = []
    body_steps = 0
    busy = 0.0
    for r in requests:
        b = batch_from_arrays([r.arrays], precision=precision)
        b.prepare(dt)
        for _ in range(r.steps):
            t = time.perf_counter()
            b.step(dt, 1)
            b.sync()
            lat.append(time.perf_counter() - t)
        busy += sum(lat[-r.steps:])
        body_steps += r.bodies * r.steps
        r.result = b.read(0)
        b.close()
    lat_a = np.array(lat) if lat else np.zeros(1)
    return {'body_steps_per_s': body_steps / busy if busy else 0.0,
            'p50_step_ms': 1e3 * float(np.percentile(lat_a, 50)),
            'p99_step_ms': 1e3 * float(np.percentile(lat_a, 99)), 'busy_s': busy}
