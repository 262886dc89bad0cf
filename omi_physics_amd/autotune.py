"""Autotuner for the per-world launches (world-to-workgroup packing).

Worlds are launched by size class (<= 256, <= 1024, <= 4096 bodies, larger):
the broadphase and the solver run one workgroup per world, and the best
workgroup size depends on how many bodies, candidates and contacts per level a
world of that class has.  For each class present in a workload, the tuner
steps a sample of that class's worlds with every candidate size, times the
``sap`` and ``solve`` kernels with HIP events, and keeps the fastest.

The result is a small JSON table (``tuned.json`` beside this file is the one
measured for the 4096-world workload on MI355X) that ``Batch`` takes as
``sap_block`` / ``solve_block``.
"""
import json
import os
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.path.join(HERE, 'tuned.json')
CANDIDATES = (64, 128, 256, 512, 1024)
CLASS_LIMITS = (256, 1024, 4096)


def size_class(n: int) -> int:
    np2 = 1
    while np2 < n:
        np2 <<= 1
    for c, lim in enumerate(CLASS_LIMITS):
        if np2 <= lim:
            return c
    return 3


def load_tuned(path: str = TUNED) -> Dict[str, List[int]]:
    """The tuned table, or defaults when none has been measured."""
    if os.path.exists(path):
        with open(path) as f:
            t = json.load(f)
        return {'sap_block': list(t['sap_block']), 'solve_block': list(t['solve_block'])}
    return {'sap_block': [64, 128, 256, 1024], 'solve_block': [64, 128, 256, 256]}


def tune(worlds: Sequence[dict], precision: str = 'fp32', dt: float = 1.0 / 60.0, sample: int = 192,
         warm_steps: int = 90, timed_steps: int = 20, candidates: Sequence[int] = CANDIDATES,
         seed: int = 0, log=print) -> Dict[str, object]:
    """Pick per-class block sizes for ``worlds`` (list of ``export_world_arrays`` dicts)."""
    from .batch import batch_from_arrays
    rng = np.random.default_rng(seed)
    by_class: Dict[int, List[int]] = {}
    for i, w in enumerate(worlds):
        by_class.setdefault(size_class(int(w['n'])), []).append(i)
    best = load_tuned()
    report = {}
    for c, idx in sorted(by_class.items()):
        pick = rng.choice(idx, size=min(sample, len(idx)), replace=False)
        arrs = [worlds[i] for i in pick]
        times = {}
        for B in candidates:
            sap = [best['sap_block'][k] if k != c else B for k in range(4)]
            sol = [best['solve_block'][k] if k != c else B for k in range(4)]
            b = batch_from_arrays(arrs, precision=precision, sap_block=sap, solve_block=sol)
            b.prepare(dt)
            b.step(dt, warm_steps)          # let the worlds settle into contact
            b.sync()
            b.set_timing(True)
            tot = {'sap': 0.0, 'solve': 0.0}
            for _ in range(timed_steps):
                b.step(dt, 1)
                b.sync()
                kt = b.kernel_times()
                tot['sap'] += kt['sap']
                tot['solve'] += kt['solve']
            b.close()
            times[B] = {k: v / timed_steps for k, v in tot.items()}
            log('class %d (%d worlds) block %4d: sap %.3f ms  solve %.3f ms'
                % (c, len(arrs), B, times[B]['sap'], times[B]['solve']))
        best['sap_block'][c] = int(min(times, key=lambda B: times[B]['sap']))
        best['solve_block'][c] = int(min(times, key=lambda B: times[B]['solve']))
        report[str(c)] = {str(B): t for B, t in times.items()}
    return {'sap_block': best['sap_block'], 'solve_block': best['solve_block'], 'precision': precision,
            'measured': report, 'when': time.strftime('%Y-%m-%d %H:%M:%S')}


def save(result: Dict[str, object], path: str = TUNED) -> None:
    with open(path, 'w') as f:
        json.dump(result, f, indent=1)


if __name__ == '__main__':
    import argparse
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    from omi_physics_amd import scenes
    ap = argparse.ArgumentParser(description='tune per-class block sizes on the 4096-world workload')
    ap.add_argument('--worlds', type=int, default=4096)
    ap.add_argument('--out', default=TUNED)
    a = ap.parse_args()
    sizes = scenes.world_sizes(a.worlds)
    worlds = [scenes.arrays(scenes.boxes_on_ground(int(sizes[w]), seed=3 * 100003 + w, sphere_fraction=0.5))
              for w in range(a.worlds)]
    res = tune(worlds, log=lambda *x: print(*x, flush=True))
    save(res, a.out)
    print(json.dumps({k: res[k] for k in ('sap_block', 'solve_block')}))
