"""Continuous batching of world rollouts on one GPU.

Requests (a world plus a number of steps) arrive over time.  The server keeps
one device batch whose body/pair/contact capacity is the memory budget; each
tick it admits waiting requests into free world slots while their bodies fit,
advances every resident world by one step in one graph-launched step, and
retires finished worlds, freeing their slots and bodies for the next
admissions.  A world's state never leaves HBM between its admission and its
retirement, and worlds admitted mid-flight share the step launches of the
worlds already running.

Host work overlaps the device: a tick queues its step and then waits only for
the previous tick's step, so the device always has the next step queued while
the host admits and retires.  Admissions are ordered on the batch stream
(``omi_batch_load_world``: one transfer on a side stream that runs during the
queued step, one copy launch before the next); a retiring world's state is
fetched into pinned memory after its last step (``omi_batch_fetch_world``)
and collected once it has arrived, and its slot is reused at once.

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
    queue_wait_s: List[float] = field(default_factory=list)
    busy_s: float = 0.0
    wall_s: float = 0.0

    def summary(self) -> Dict[str, float]:
        lat = np.array(self.step_latency_s) if self.step_latency_s else np.zeros(1)
        qw = np.array(self.queue_wait_s) if self.queue_wait_s else np.zeros(1)
        bs = float(np.sum(self.active_bodies))
        return {'ticks': len(self.step_latency_s),
                'body_steps': bs,
                'body_steps_per_s': bs / self.wall_s if self.wall_s > 0 else 0.0,
                'p50_step_ms': 1e3 * float(np.percentile(lat, 50)),
                'p99_step_ms': 1e3 * float(np.percentile(lat, 99)),
                'queue_wait_p50_ms': 1e3 * float(np.percentile(qw, 50)),
                'queue_wait_p99_ms': 1e3 * float(np.percentile(qw, 99)),
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
        self.fetching: List[Any] = []
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
                self.batch.load(slot, req.arrays)      # ordered after the queued step; no wait
            except RuntimeError as e:
                if 'capacity' in str(e):
                    break                  # fragmented arena: wait for retirements
                raise
            self.waiting.popleft()
            self.free_slots.popleft()
            req.slot = slot
            req.admitted_at = now
            self.stats.queue_wait_s.append(max(0.0, now - req.arrival))
            self.resident[slot] = req
            self.bodies_in_use += req.bodies
            admitted += 1
        return admitted

    def _retire(self) -> None:
        """Worlds whose last step is queued: fetch their state after it, free the slot."""
        for slot in [s for s, r in self.resident.items() if r.done_steps >= r.steps]:
            req = self.resident.pop(slot)
            self.fetching.append((req, self.batch.fetch(slot)))
            self.batch.remove(slot)
            self.free_slots.append(slot)
            self.bodies_in_use -= req.bodies

    def _collect(self, now: float, wait: bool = False) -> None:
        keep = []
        for req, ticket in self.fetching:
            if wait or self.batch.fetch_ready(ticket):
                req.result = self.batch.fetched(ticket)
                req.finished_at = now
                self.finished.append(req)
            else:
                keep.append((req, ticket))
        self.fetching = keep

    def run(self, requests: List[Request], realtime: bool = True) -> ServeStats:
        """Serve ``requests`` (sorted by arrival).  With ``realtime`` the clock is
        wall time since start, with idle gaps skipped ahead rather than slept.

        A tick admits arrivals, queues one step of every resident world and a
        marker after it, retires the worlds that step finishes, then waits for
        the previous tick's marker: the step latency is the time from one
        step's end (or its launch, if later) to the next one's."""
        for r in sorted(requests, key=lambda r: r.arrival):
            self.submit(r)
        t0 = time.perf_counter()
        skew = 0.0
        self.batch.prepare(self.dt)
        inflight = None                    # (marker, launch time, bodies stepped)
        last_end = t0
        while self.waiting or self.resident or inflight is not None or self.fetching:
            now = (time.perf_counter() - t0) + skew if realtime else float('inf')
            self._admit(now)
            queued = None
            if self.resident:
                t_launch = time.perf_counter()
                self.batch.step(self.dt, 1)
                queued = (self.batch.mark(), t_launch, self.bodies_in_use)
                for r in self.resident.values():
                    r.done_steps += 1
                self._retire()
            if inflight is not None:
                mk, t_launch, bodies = inflight
                self.batch.mark_wait(mk)
                t_end = time.perf_counter()
                self.stats.step_latency_s.append(t_end - max(t_launch, last_end))
                self.stats.active_bodies.append(bodies)
                last_end = t_end
            inflight = queued
            self._collect(now)
            if inflight is None and not self.resident:
                if self.fetching:
                    self._collect(now, wait=True)
                elif self.waiting:
                    skew += max(0.0, self.waiting[0].arrival - now)   # idle: jump to the next arrival
                    last_end = time.perf_counter()
        self.batch.sync()
        self.stats.wall_s = time.perf_counter() - t0
        self.stats.busy_s = float(np.sum(self.stats.step_latency_s))
        return self.stats


def run_plain(requests: List[Request], precision: str = 'fp32', dt: float = 1.0 / 60.0) -> Dict[str, float]:
    """The plain form: each request alone on the GPU, one after another, each
    step waited for.  ``body_steps_per_s`` is over the whole run (loads and
    readbacks included, as the server's is); ``device_body_steps_per_s`` over
    the steps' own time."""
    from .batch import batch_from_arrays
    lat: List[float] = []
    body_steps = 0
    busy = 0.0
    t0 = time.perf_counter()
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
    wall = time.perf_counter() - t0
    lat_a = np.array(lat) if lat else np.zeros(1)
    return {'body_steps_per_s': body_steps / wall if wall else 0.0,
            'device_body_steps_per_s': body_steps / busy if busy else 0.0,
            'p50_step_ms': 1e3 * float(np.percentile(lat_a, 50)),
            'p99_step_ms': 1e3 * float(np.percentile(lat_a, 99)), 'busy_s': busy, 'wall_s': wall}
