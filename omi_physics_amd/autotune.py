"""Autotuner for the per-world launches (world-to-workgroup packing).

Worlds are launched by size class (<= 256, <= 1024, <= 4096 bodies, larger):
the broadphase sort and the solver run one workgroup per world, the classes
side by side.  The best workgroup size of a class depends on its worlds'
bodies, contacts per level and on how many of them share a CU, so the tuner
times the whole step on the whole workload: after the worlds have settled
into contact it tries every candidate size for one class and kernel at a
time (coordinate descent), interleaving the candidates over a few rounds so
the slowly changing contact state weighs on all of them alike, and keeps the
fastest.

The result is a small JSON table per precision (``tuned.json`` beside this
file holds the ones measured for the 4096-world workload on MI355X) that
``Batch`` takes as ``sap_block`` / ``solve_block``.  Classes absent from the
tuning workload keep their table entry: the class of worlds over 4096 bodies
was timed on the 9101-body pyramids-and-debris world (FP32 solver 512
threads, FP64 256).  FP64 solver blocks stop
at 256 threads: the FP64 sweep needs more registers per thread than a larger
block leaves.
"""
import json
import os
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.path.join(HERE, 'tuned.json')
CANDIDATES = (64, 128, 256, 512, 1024)
FP64_SOLVE_BLOCK_MAX = 256
CLASS_LIMITS = (256, 1024, 4096)
DEFAULTS = {'fp32': {'sap_block': [64, 128, 1024, 1024], 'solve_block': [64, 128, 256, 512]},
            'fp64': {'sap_block': [64, 128, 1024, 1024], 'solve_block': [64, 128, 256, 256]}}


def _prec(precision) -> str:
    return 'fp64' if precision in ('fp64', 64) else 'fp32'


def size_class(n: int) -> int:
    np2 = 1
    while np2 < n:
        np2 <<= 1
    for c, lim in enumerate(CLASS_LIMITS):
        if np2 <= lim:
            return c
    return 3


def load_tuned(precision='fp32', path: str = TUNED) -> Dict[str, List[int]]:
    """The tuned table of ``precision``, or the defaults when none has been measured."""
    p = _prec(precision)
    t = {}
    if os.path.exists(path):
        with open(path) as f:
            t = json.load(f).get(p, {})
    d = DEFAULTS[p]
    return {k: [int(x) for x in t.get(k, d[k])] for k in ('sap_block', 'solve_block')}


# Size classes 1-3 holding at most this many worlds leave most CUs idle, so
# their worlds take the largest blocks instead of the crowded-batch table
# (one 1001-body world: FP64 0.740 -> 0.715 ms/step with a 256-thread solver
# and 1024-thread sort; FP32 0.580 -> 0.533 ms/step with a 512-thread solver).
SPARSE_WORLDS = 16
SPARSE_BLOCKS = {'fp64': {'sap_block': 1024, 'solve_block': 256},
                 'fp32': {'sap_block': 1024, 'solve_block': 512}}


def blocks_for(sizes: Sequence[int], precision: str = 'fp32') -> Dict[str, List[int]]:
    """Block sizes for a batch of worlds with these body counts: the tuned
    table, with the sparse blocks for classes 1-3 holding few worlds."""
    p = _prec(precision)
    t = load_tuned(p)
    counts = [0, 0, 0, 0]
    for n in sizes:
        counts[size_class(int(n))] += 1
    for c in (1, 2, 3):
        if 0 < counts[c] <= SPARSE_WORLDS:
            for k in ('sap_block', 'solve_block'):
                t[k][c] = max(t[k][c], SPARSE_BLOCKS[p][k])
    return t


def tune(worlds: Sequence[dict], precision: str = 'fp32', dt: float = 1.0 / 60.0, warm_steps: int = 120,
         timed_steps: int = 6, rounds: int = 3, candidates: Sequence[int] = CANDIDATES,
         log=print) -> Dict[str, object]:
    """Pick per-class block sizes for ``worlds`` (list of ``export_world_arrays`` dicts)."""
    from .batch import batch_from_arrays
    classes = sorted({size_class(int(w['n'])) for w in worlds})
    best = load_tuned(precision)
    b = batch_from_arrays(worlds, precision=precision, sap_block=best['sap_block'], solve_block=best['solve_block'])
    b.prepare(dt)
    b.step(dt, warm_steps)          # let the worlds settle into contact
    b.sync()
    report = {}
    for c in classes:
        for kind in ('sap_block', 'solve_block'):
            cands = [B for B in candidates
                     if not (kind == 'solve_block' and _prec(precision) == 'fp64' and B > FP64_SOLVE_BLOCK_MAX)]
            times = {B: 0.0 for B in cands}
            for _ in range(rounds):
                for B in cands:
                    cfg = {k: list(v) for k, v in best.items()}
                    cfg[kind][c] = B
                    b.set_blocks(cfg['sap_block'], cfg['solve_block'])
                    b.prepare(dt)
                    b.sync()
                    t = time.perf_counter()
                    b.step(dt, timed_steps)
                    b.sync()
                    times[B] += (time.perf_counter() - t) / (rounds * timed_steps)
            pick = int(min(times, key=times.get))
            best[kind][c] = pick
            report['%s[%d]' % (kind, c)] = {str(B): round(v * 1e3, 4) for B, v in times.items()}
            log('class %d %s: %s ms/step -> %d' % (c, kind, ' '.join('%d:%.3f' % (B, v * 1e3)
                                                                      for B, v in times.items()), pick))
    b.close()
    return {'sap_block': best['sap_block'], 'solve_block': best['solve_block'], 'precision': _prec(precision),
            'workload_worlds': len(worlds), 'measured_ms_per_step': report,
            'when': time.strftime('%Y-%m-%d %H:%M:%S')}


def save(result: Dict[str, object], path: str = TUNED) -> None:
    """Store ``result`` as its precision's table, keeping the other precision's."""
    tables = {}
    if os.path.exists(path):
        with open(path) as f:
            tables = {k: v for k, v in json.load(f).items() if k in DEFAULTS}
    tables[_prec(result['precision'])] = result
    with open(path, 'w') as f:
        json.dump(tables, f, indent=1)


if __name__ == '__main__':
    import argparse
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    from omi_physics_amd import scenes
    ap = argparse.ArgumentParser(description='tune per-class block sizes on the 4096-world workload')
    ap.add_argument('--worlds', type=int, default=4096)
    ap.add_argument('--out', default=TUNED)
    ap.add_argument('--precision', choices=('fp32', 'fp64', 'both'), default='both')
    a = ap.parse_args()
    sizes = scenes.world_sizes(a.worlds)
    worlds = [scenes.arrays(scenes.boxes_on_ground(int(sizes[w]), seed=3 * 100003 + w, sphere_fraction=0.5))
              for w in range(a.worlds)]
    for prec in (('fp32', 'fp64') if a.precision == 'both' else (a.precision,)):
        res = tune(worlds, precision=prec, log=lambda *x: print(prec, *x, flush=True))
        save(res, a.out)
        print(json.dumps({'precision': prec, 'sap_block': res['sap_block'], 'solve_block': res['solve_block']}),
              flush=True)
