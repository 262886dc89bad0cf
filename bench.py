"""Benchmark: body-steps/s of the batched contact pipeline on one MI355X.

Workload (the headline): 4096 independent worlds, each a static ground box
(half extents 10 x 0.5 x 10) plus n_w dynamic bodies, n_w =
clip(round(lognormal(ln 64, 1.0)), 8, 2048) (seed 3, 436778 bodies), a 50/50
mix of unit boxes and r = 0.5 spheres of mass 1 dropped from y in [2, 50] over
x, z in [-10, 10]; mu = 0.5, e = 0, dt = 1/60, 10 solver iterations, sleeping
on; a 300-step rollout.  All worlds are one batch stepped by one graph launch
per step, FP64 (bit-identical to the oracle).

Timed window: the last ``--steps`` K steps of the rollout.  The batch is
advanced untimed through steps 1..300-K-W, then W warm-up steps, then the K
steps are launched back to back and timed as one job (a device
synchronisation on both sides): ``value`` and ``ms_per_step``.  A second
batch runs the same rollout with every step synchronised, for the per-step
latency (``p50_step_ms``, ``p99_step_ms``, ``ms_per_step_synced``) and the
counters of every step, so the line states the window's contacts per step
next to the whole rollout's (its results equal the first batch's bit for
bit).  With the driver's K = 20, W = 5 the window is steps 281..300, after the
piles have settled (contacts per step above the rollout's mean).

Printed: one JSON line (details on stderr).  Beside ``value``: the batched FP32
form on the same window, per-stage device time and the roofline of the
dominant kernel and of the whole step from the same window, the plain
one-world-at-a-time GPU form, the CPU baseline (the oracle on every world over
the same window), parity of all 4096 worlds with the oracle at the window's
end, FP32 error curves against FP64 beside the FP64 one-ulp control, a
continuous-batching serving run, and the single-world workloads.
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = 'body-steps/s at 1 GPU (batched vs plain) + p99 step latency'
DT = 1.0 / 60.0
ROLLOUT = 300
HBM_PEAK_GBS = 8000.0
# worker threads for the CPU legs: the host cores given to one GPU
CPU_THREADS = min(len(os.sched_getaffinity(0)), 16)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_worlds(n_worlds, seed):
    from omi_physics_amd import scenes
    sizes = scenes.world_sizes(n_worlds, seed=seed)
    t = time.perf_counter()
    arrays = [scenes.arrays(scenes.boxes_on_ground(int(sizes[w]), seed=seed * 100003 + w, sphere_fraction=0.5))
              for w in range(n_worlds)]
    log('scene: %d worlds, %d bodies (%.1f s)' % (n_worlds, sum(a['n'] for a in arrays), time.perf_counter() - t))
    return arrays


# ----------------------------------------------------------- algorithmic bytes

def step_bytes(c, bodies_in_contact, r):
    """HBM bytes one whole step must move (batched form, solver float r bytes):
    376 per body (refit 120, integrate 180, sleep 44, sort keys 32), 32 per
    x-window candidate, 152 per solid pair, 16 per manifold, 80 per contact, 144
    per body in a contacting island; the FP64 form adds 24 per body and 16 per
    contact (FP64 velocities and impulses)."""
    N, X, P, M, C = c['bodies'], c['x_candidates'], c['solid_pairs'], c['manifolds'], c['contacts']
    extra = (24 * N + 16 * C) if r == 8 else 0
    return 376 * N + 32 * X + 152 * P + 16 * M + 80 * C + 144 * bodies_in_contact + extra


def stage_bytes(kernel, c, r):
    """HBM bytes one launch of a stage must move, from the step's counters:
      pre     350 + 18r per body (pose, velocity, parameters in; previous pose,
              velocity, AABB, rotation and world inverse inertia out)
      sap     16 per body (lo.x, hi.x) + 32 per x-window candidate (y/z bounds)
              + 8 per solid pair written
      narrow  106 + 18r per solid pair (indices, kinds, poses, rotations, half
              extents) + 24 + r per contact point written
      solve   32 + 7r per contact (indices, point, normal, depth, impulses)
              + 152 + 9r per body touched, taken as min(bodies, 2 x manifolds)
    """
    N, X, P, M, C = c['bodies'], c['x_candidates'], c['solid_pairs'], c['manifolds'], c['contacts']
    if kernel == 'pre':
        return (350 + 18 * r) * N
    if kernel == 'sap':
        return 16 * N + 32 * X + 8 * P
    if kernel == 'narrow':
        return (106 + 18 * r) * P + (24 + r) * C
    if kernel == 'solve':
        return (32 + 7 * r) * C + (152 + 9 * r) * min(N, 2 * M)
    return 0


def solve_top_bytes(stats, r):
    """Algorithmic bytes of the solver launch of the largest size class present:
    per contact its 224-byte row in and three impulses (3r) out, per manifold its
    16-byte record, per body in contact (min(bodies, 2 x manifolds) per world) FP64
    velocity in and out (96), FP64 position in and out (48) and world inverse
    inertia (12r); with more than one class present the launch also writes its
    contacts' rows (224 out and 60 of contact data in per contact, as k_prepare)."""
    top = int(stats['size_class'].max())
    sel = stats['size_class'] == top
    k = int(stats['contacts'][sel].sum())
    m = int(stats['manifolds'][sel].sum())
    nd = int(np.minimum(stats['bodies'][sel], 2 * stats['manifolds'][sel]).sum())
    own_rows = len(set(stats['size_class'][stats['size_class'] >= 0].tolist())) > 1
    return (k * (224 + 3 * r) + 16 * m + nd * (96 + 48 + 12 * r) + (k * (224 + 60) if own_rows else 0),
            top, int(sel.sum()))


def bodies_in_contact(stats):
    return int(np.minimum(stats['bodies'], 2 * stats['manifolds']).sum())


# ------------------------------------------------------------------ rollouts

def window_bounds(steps, warmup):
    """(first step of the warm-up, first timed step, last timed step, rollout length)."""
    first = max(1, ROLLOUT - steps - warmup + 1)
    t0 = first + warmup
    return first, t0, t0 + steps - 1, max(ROLLOUT, t0 + steps - 1)


def profile_summary(path, last):
    """Per-kernel figures of a rocprofv3 CSV output directory over each
    kernel's last ``last`` dispatches (with ``--only-headline`` and a kernel
    launched once per step: the timed window): from a kernel trace the mean
    and total duration, from a counter collection the mean of each counter per
    dispatch.  Returns markdown table lines."""
    import csv
    import glob
    import re
    lines = []
    for f in sorted(glob.glob(os.path.join(path, '**', '*.csv'), recursive=True)):
        with open(f) as fh:
            rows = list(csv.DictReader(fh))
        if not rows or 'Kernel_Name' not in rows[0]:
            continue
        short = lambda n: re.sub(r'\(.*', '', n).strip()
        if 'Counter_Name' in rows[0]:
            per = {}
            for r in rows:
                key = (short(r['Kernel_Name']), r['Counter_Name'])
                d = per.setdefault(key, {})
                d[int(r['Dispatch_Id'])] = d.get(int(r['Dispatch_Id']), 0.0) + float(r['Counter_Value'])
            lines += ['', '%s: mean per dispatch over the last %d dispatches per kernel' % (os.path.basename(f), last),
                      '', '| kernel | counter | dispatches | mean |', '|---|---|---|---|']
            tab = []
            for (k, c), d in per.items():
                v = [d[i] for i in sorted(d)][-last:]
                tab.append((k, c, len(v), sum(v) / len(v)))
            for k, c, nd, m in sorted(tab, key=lambda t: -t[3]):
                lines.append('| `%s` | %s | %d | %.1f |' % (k, c, nd, m))
        elif 'Start_Timestamp' in rows[0]:
            per = {}
            for r in sorted(rows, key=lambda r: int(r['Start_Timestamp'])):
                per.setdefault(short(r['Kernel_Name']), []).append(
                    (int(r['End_Timestamp']) - int(r['Start_Timestamp'])) * 1e-3)
            lines += ['', '%s: last %d dispatches per kernel' % (os.path.basename(f), last), '',
                      '| kernel | dispatches | mean us | total us |', '|---|---|---|---|']
            tab = sorted(((k, v[-last:]) for k, v in per.items()), key=lambda kv: -sum(kv[1]))
            for k, v in tab:
                lines.append('| `%s` | %d | %.1f | %.1f |' % (k, len(v), sum(v) / len(v), sum(v)))
    return lines


def rollout(arrays, precision, steps, warmup, timing=False, record=True, chained=False):
    """Step one batch of ``arrays`` through the rollout; time the window.

    Returns the batch (state at the window's end) and a dict: window wall time,
    per-step latencies of the window and of the rollout, counters of every step,
    and with ``timing`` the per-stage device time and algorithmic bytes of the
    window (its steps launched eagerly with HIP events between the stages; every
    other step is one graph launch).  With ``chained`` the window's steps are
    launched back to back, one device synchronisation before and one after
    (a job's throughput): no per-step latency or counters inside the window."""
    from omi_physics_amd.batch import batch_from_arrays
    first, t0, t1, total = window_bounds(steps, warmup)
    tl = time.perf_counter()
    b = batch_from_arrays(arrays, precision=precision)
    log('%s batch loaded: %.1f GB device, %.1f s' % (precision, b.device_bytes() / 1e9, time.perf_counter() - tl))
    b.prepare(DT)
    b.sync()
    r = 4 if precision == 'fp32' else 8
    lat, cnt = [], []
    kt, kb = {}, {}
    wall = 0.0
    for s in range(1, t1 + 1):
        if s == t0:
            if timing:
                b.set_timing(True)
            b.sync()
            tw = time.perf_counter()
            if chained:
                for _ in range(t0, t1 + 1):
                    b.step(DT, 1)
                b.sync()
                wall = time.perf_counter() - tw
                break
        t = time.perf_counter()
        b.step(DT, 1)
        b.sync()
        lat.append(time.perf_counter() - t)
        if record:
            cnt.append(b.counters())
        if timing and s >= t0:
            for k, v in b.kernel_times().items():
                kt[k] = kt.get(k, 0.0) + v
            c = cnt[-1]
            st = b.world_stats()
            for k in ('pre', 'sap', 'narrow', 'solve'):
                kb[k] = kb.get(k, 0) + stage_bytes(k, c, r)
            tb, top, ntop = solve_top_bytes(st, r)
            kb['solve_top'] = kb.get('solve_top', 0) + tb
            kb['top_class'] = (top, ntop)
            kb['step'] = kb.get('step', 0) + step_bytes(c, bodies_in_contact(st), r)
        if s == t1:
            wall = time.perf_counter() - tw
        if s % 50 == 0:
            log('  %s step %d/%d  %.2f ms' % (precision, s, total, 1e3 * lat[-1]))
    if timing:
        b.set_timing(False)
    return b, dict(wall=wall, lat=np.array(lat), window_lat=np.array(lat[t0 - 1:t1]), counters=cnt,
                   kernel_ms=kt, kernel_bytes=kb, bounds=(first, t0, t1, total))


def finish_rollout(b, info, steps_done):
    """Step the batch untimed to the end of the rollout, recording counters and latency."""
    first, t0, t1, total = info['bounds']
    for s in range(steps_done + 1, total + 1):
        t = time.perf_counter()
        b.step(DT, 1)
        b.sync()
        info['lat'] = np.append(info['lat'], time.perf_counter() - t)
        info['counters'].append(b.counters())


def mean_counter(cs, key):
    return float(np.mean([c[key] for c in cs])) if cs else 0.0


# ----------------------------------------------------------------- CPU legs

def oracle_window(arrays, steps, warmup, threads):
    """The oracle on every world through the same window: each world advanced
    untimed to the window, then the window's steps timed over all worlds."""
    import oracle
    first, t0, t1, _ = window_bounds(steps, warmup)
    worlds = [oracle.OracleWorld(a) for a in arrays]

    def advance(ow):
        ow.step(DT, t0 - 1)

    def window(ow):
        ow.step(DT, steps)
    ta = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(advance, worlds))
    tw = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(window, worlds))
    wall = time.perf_counter() - tw
    log('oracle: %d worlds advanced %d steps in %.1f s, window of %d steps in %.2f s on %d threads'
        % (len(worlds), t0 - 1, tw - ta, steps, wall, threads))
    return worlds, wall


def error_curves(arrays, checkpoints, threads):
    """Position error against the FP64 oracle at each checkpoint step (median,
    90th percentile and largest over worlds of each world's largest error): the
    batched FP32 form, and the FP64 oracle started from positions moved by one ulp
    (how far the piles carry a last-bit change on their own)."""
    import oracle
    from omi_physics_amd.batch import batch_from_arrays

    def nudged(a):
        a = dict(a)
        p = a['position'].copy()
        dyn = a['motion_type'] == 2
        p[dyn] = np.nextafter(p[dyn], np.inf)
        a['position'] = p
        return a

    def trace(a):
        ow = oracle.OracleWorld(a)
        out, s = [], 0
        for c in checkpoints:
            ow.step(DT, c - s)
            s = c
            out.append(ow.state()['position'])
        return out
    with ThreadPoolExecutor(max_workers=threads) as ex:
        ref = list(ex.map(trace, arrays))
        ulp = list(ex.map(trace, [nudged(a) for a in arrays]))
    b = batch_from_arrays(arrays, precision='fp32')
    f32, s = [], 0
    for c in checkpoints:
        b.step(DT, c - s)
        b.sync()
        s = c
        f32.append([p for p, _, _, _ in b.snapshot()])
    b.close()

    def stats(errs):
        e = np.asarray(errs)
        return {'median_m': float(np.median(e)), 'p90_m': float(np.percentile(e, 90)), 'max_m': float(e.max()),
                'worlds_within_1e-4': float(np.mean(e <= 1e-4))}
    curve = {}
    for k, c in enumerate(checkpoints):
        e32 = [np.abs(f32[k][w] - ref[w][k]).max() for w in range(len(arrays))]
        eul = [np.abs(ulp[w][k] - ref[w][k]).max() for w in range(len(arrays))]
        curve[str(c)] = {'fp32': stats(e32), 'fp64_1ulp': stats(eul)}
    return curve


# -------------------------------------------------------------- other forms

def run_plain(arrays, steps, warmup, precision):
    """The plain form: each sampled world alone on the GPU, one world at a time,
    over the same window of its rollout."""
    from omi_physics_amd.batch import batch_from_arrays
    first, t0, t1, _ = window_bounds(steps, warmup)
    body_steps, total, lat = 0, 0.0, []
    for a in arrays:
        b = batch_from_arrays([a], precision=precision)
        b.prepare(DT)
        b.step(DT, t0 - 1)
        b.sync()
        ts = time.perf_counter()
        for _ in range(steps):
            t = time.perf_counter()
            b.step(DT, 1)
            b.sync()
            lat.append(time.perf_counter() - t)
        total += time.perf_counter() - ts
        body_steps += a['n'] * steps
        b.close()
    return body_steps / total, np.array(lat)


def serving_run(rates, n_requests, precision, seed=5):
    """Continuous batching of Poisson-arriving rollouts (8..4096 bodies, 60 steps)
    at each arrival rate in ``rates`` (the same worlds: one seed, arrival times
    scaled), against the same requests served one at a time."""
    from omi_physics_amd import scenes
    from omi_physics_amd.serve import Request, RolloutServer, run_plain as serve_plain

    base = scenes.serving_requests(n_requests, rates[0], seed=seed)
    arrays = [scenes.arrays(scenes.boxes_on_ground(r['n'], seed=r['seed'], sphere_fraction=0.5)) for r in base]
    out = {'requests': n_requests, 'precision': precision, 'steps_per_request': base[0]['steps'],
           'bodies_per_request': [int(min(r['n'] for r in base)), int(max(r['n'] for r in base))],
           'rates': 'body-steps over the run\'s wall time (admissions and retirements included; the clock skips '
                    'idle gaps between arrivals)', 'sweep': []}
    for rate in rates:
        spec = scenes.serving_requests(n_requests, rate, seed=seed)
        reqs = [Request(id=i, arrays=arrays[i], steps=r['steps'], arrival=r['arrival']) for i, r in enumerate(spec)]
        srv = RolloutServer(max_worlds=1024, max_bodies=1 << 19, precision=precision)
        st = srv.run(reqs).summary()
        srv.close()
        wait = [r.admitted_at - r.arrival for r in reqs]
        out['sweep'].append({'arrival_rate_per_s': rate, 'body_steps_per_s': st['body_steps_per_s'],
                             'p50_step_ms': st['p50_step_ms'], 'p99_step_ms': st['p99_step_ms'],
                             'queue_wait_p50_ms': 1e3 * float(np.percentile(wait, 50)),
                             'queue_wait_p99_ms': 1e3 * float(np.percentile(wait, 99)),
                             'mean_active_bodies': st['mean_active_bodies'], 'ticks': st['ticks'],
                             'wall_s': st['wall_s']})
        log('serving %s at %g/s: %s' % (precision, rate, out['sweep'][-1]))
    plain = serve_plain([Request(id=i, arrays=arrays[i], steps=base[i]['steps']) for i in range(24)],
                        precision=precision)
    out['plain'] = {'body_steps_per_s': plain['body_steps_per_s'],
                    'device_body_steps_per_s': plain['device_body_steps_per_s'],
                    'p50_step_ms': plain['p50_step_ms'], 'p99_step_ms': plain['p99_step_ms'], 'requests': 24}
    top = max(out['sweep'], key=lambda e: e['body_steps_per_s'])
    out.update({k: top[k] for k in ('arrival_rate_per_s', 'body_steps_per_s', 'p50_step_ms', 'p99_step_ms',
                                     'queue_wait_p99_ms')})
    out['over_plain'] = top['body_steps_per_s'] / max(plain['body_steps_per_s'], 1e-9)
    return out


def single_world(name, arrays, precision, steps, warmup, cpu):
    """One world stepped alone on the GPU (the plain form) over its 300-step
    rollout, the last ``steps`` steps timed; the oracle on one core over the same
    rollout, its state compared at the end."""
    import oracle
    n = int(arrays['n'])
    b, info = rollout([arrays], precision, steps, warmup, record=True)
    first, t0, t1, total = info['bounds']
    finish_rollout(b, info, t1)
    pos = b.read(0)['position']
    b.close()
    out = {'workload': name, 'bodies': n, 'precision': precision, 'window': [t0, t1],
           'value': n * steps / info['wall'], 'ms_per_step': 1e3 * info['wall'] / steps,
           'rollout_ms_per_step': 1e3 * float(info['lat'].mean()),
           'p50_step_ms': 1e3 * float(np.percentile(info['window_lat'], 50)),
           'p99_step_ms': 1e3 * float(np.percentile(info['window_lat'], 99)),
           'contacts_per_step_window': mean_counter(info['counters'][t0 - 1:t1], 'contacts'),
           'contacts_per_step_rollout': mean_counter(info['counters'], 'contacts')}
    if cpu:
        ow = oracle.OracleWorld(arrays)
        tc = time.perf_counter()
        ow.step(DT, total)
        cw = time.perf_counter() - tc
        out['cpu_baseline'] = {'value': n * total / cw, 'unit': 'body-steps/s', 'cores': 1, 'kind': 'port',
                               'sample': 'the whole world, all %d steps, one core (%.1f s)' % (total, cw)}
        d = np.abs(ow.state()['position'] - pos).max()
        out['max_pos_err_vs_fp64_oracle_m'] = float(d)
    log('single world %s: %s' % (name, {k: v for k, v in out.items() if k != 'workload'}))
    return out


SINGLE = ('boxes1000', 'pyramids', 'convex_chains')


def single_arrays(which):
    """(description, arrays) of a single-world workload."""
    from omi_physics_amd import scenes, PhysicsWorld, model
    if which == 'boxes1000':
        return ('1000 unit boxes on a ground box (20 x 20 m), one world',
                scenes.arrays(scenes.boxes_on_ground(1000, seed=1)))
    if which == 'pyramids':
        return ('20 ten-level box pyramids + 8000 box/sphere debris, one world',
                scenes.arrays(scenes.pyramids_and_debris()))
    t = time.perf_counter()
    cvx = scenes.build(scenes.convex_and_chains(), PhysicsWorld, model).export_arrays()
    log('convex-and-chains world built in %.1f s' % (time.perf_counter() - t))
    return '2000 cooked convex hulls (16-32 points) + 20 hinge chains of 50 links, one world', cvx


def single_worlds(steps, warmup, cpu):
    res = {}
    name, arrays = single_arrays('boxes1000')
    res['boxes1000_fp64'] = single_world(name, arrays, 'fp64', steps, warmup, cpu)
    for which in ('pyramids', 'convex_chains'):
        name, arrays = single_arrays(which)
        for prec in ('fp32', 'fp64'):
            res['%s_%s' % (which, prec)] = single_world(name, arrays, prec, steps, warmup, cpu and prec == 'fp64')
    return res


# ---------------------------------------------------------------------- main

def copy_bench(mib):
    """A plain device-to-device copy of known size through the HIP runtime: the
    achievable HBM rate and the calibration of the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL(os.path.join(os.environ.get('ROCM_PATH', '/opt/rocm'), 'lib', 'libamdhip64.so'))
    nbytes = ctypes.c_size_t(mib << 20)
    x, y = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(x), nbytes) or hip.hipMalloc(ctypes.byref(y), nbytes):
        raise RuntimeError('hipMalloc failed')
    hip.hipMemset(x, 1, nbytes)
    for _ in range(3):
        hip.hipMemcpyAsync(y, x, nbytes, 3, None)
    hip.hipDeviceSynchronize()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    reps = 20
    hip.hipEventRecord(e0, None)
    for _ in range(reps):
        hip.hipMemcpyAsync(y, x, nbytes, 3, None)
    hip.hipEventRecord(e1, None)
    hip.hipEventSynchronize(e1)
    ms = ctypes.c_float()
    hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
    per = ms.value / reps
    print(json.dumps({'workload': 'copy', 'bytes_read_per_copy': nbytes.value, 'bytes_written_per_copy':
                      nbytes.value, 'ms_per_copy': per, 'GBps': 2 * nbytes.value / (per * 1e-3) / 1e9}), flush=True)
    hip.hipFree(x)
    hip.hipFree(y)


def traffic_for(window, workload):
    """Measured HBM bytes per launch of the dominant kernel from the committed
    rocprof summary, if it was taken on this window of this workload."""
    tf = os.path.join(HERE, 'profiles', 'traffic.json')
    if not os.path.exists(tf):
        return None, None
    with open(tf) as f:
        t = json.load(f)
    if t.get('window') != list(window) or t.get('workload') != workload:
        return None, None
    return t.get('k_solve_top_fp64_bytes_per_launch'), t.get('command')


def solve_roofline(ti, K, window):
    """The roofline of the largest size class's solver launch from an event-timed pass."""
    kt, kb = ti['kernel_ms'], ti['kernel_bytes']
    top, ntop = kb['top_class']
    avg_ms = kt['solve_top'] / K
    per_launch = kb['solve_top'] / K
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    traffic, tcmd = traffic_for(window, 'batched4096')
    frac = achieved / HBM_PEAK_GBS
    return {
        'kernel': 'k_solve<double> size class %d (%d worlds, the largest; alone on its stream)' % (top, ntop),
        'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': frac,
        'traffic': traffic, 'traffic_command': tcmd, 'avg_ms': avg_ms, 'bytes_per_launch': per_launch,
        'limiter': ('latency: %.2f %% of the HBM roof; the launch is one dependency chain of solver levels '
                    '(occupancy and wave stalls: profiles/sq_stalls_solver_fp64.md)' % (100 * frac)
                    if frac < 0.05 else 'hbm')}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20, help='timed steps: the last steps of the 300-step rollout')
    ap.add_argument('--warmup', type=int, default=5, help='untimed steps right before the timed window')
    ap.add_argument('--worlds', type=int, default=4096)
    ap.add_argument('--seed', type=int, default=3)
    ap.add_argument('--workload', default='batched4096',
                    choices=['batched4096', 'copy', 'boxes1000', 'pyramids', 'convex_chains'],
                    help='the headline batch, a plain copy (counter calibration), or one single-world workload '
                         '(FP64, its 300-step rollout, the last --steps steps timed)')
    ap.add_argument('--copy-mib', type=int, default=1024, help='buffer size of the copy workload')
    ap.add_argument('--plain-sample', type=int, default=48, help='worlds stepped alone for the plain form')
    ap.add_argument('--curve-sample', type=int, default=192, help='worlds of the FP32 error curves')
    ap.add_argument('--only-headline', action='store_true', help='the FP64 window and nothing else (profiling)')
    ap.add_argument('--only-timing', action='store_true',
                    help='the event-timed pass of the FP64 window and its roofline, nothing else (profiling)')
    ap.add_argument('--no-timing', action='store_true', help='skip the event-timed pass of the window')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-extras', action='store_true', help='skip serving and the single-world workloads')
    ap.add_argument('--profile-summary', metavar='DIR',
                    help='print per-kernel figures of a rocprofv3 CSV directory over the last --steps dispatches')
    args = ap.parse_args()
    if args.profile_summary:
        print('\n'.join(profile_summary(args.profile_summary, args.steps)))
        return
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU: every GPU box here has one.  The independent worlds would shard '
              'across GPUs by world id (greedy bin-packing by body count, no collective on the step path).',
              file=sys.stderr)
        sys.exit(2)
    from omi_physics_amd import _ffi
    if _ffi.load().omi_device_count() < 1:
        print('no HIP device visible', file=sys.stderr)
        sys.exit(1)
    if args.workload == 'copy':
        copy_bench(args.copy_mib)
        return
    if args.workload in SINGLE:
        name, arrays = single_arrays(args.workload)
        print(json.dumps(single_world(name, arrays, 'fp64', args.steps, args.warmup, not args.no_cpu)), flush=True)
        return

    arrays = make_worlds(args.worlds, args.seed)
    n_bodies = int(sum(a['n'] for a in arrays))
    first, t0, t1, total = window_bounds(args.steps, args.warmup)
    if args.only_timing:
        bt, ti = rollout(arrays, 'fp64', args.steps, args.warmup, timing=True)
        bt.close()
        print(json.dumps({'roofline': solve_roofline(ti, args.steps, (t0, t1)),
                          'kernels_ms_per_step': {k: v / args.steps for k, v in ti['kernel_ms'].items()}}), flush=True)
        return
    workload = ('%d independent box/sphere worlds, lognormal(ln 64, 1.0) bodies clipped [8, 2048], batched in one '
                'graph-captured step, FP64 (bit-identical to the oracle); timed window: steps %d..%d of the %d-step '
                'rollout' % (args.worlds, t0, t1, total))

    # headline: the batched FP64 form, graph-launched, the window's steps
    # launched back to back and timed as one job
    b, hinfo = rollout(arrays, 'fp64', args.steps, args.warmup, chained=True)
    wall = hinfo['wall']
    value = n_bodies * args.steps / wall
    snap64 = [p for p, _, _, _ in b.snapshot()]
    b.close()
    # the same window again with every step synchronised: per-step latency and
    # the step counters
    b, info = rollout(arrays, 'fp64', args.steps, args.warmup)
    same_synced = all(np.array_equal(p, q) for p, (q, _, _, _) in zip(snap64, b.snapshot()))
    finish_rollout(b, info, t1)
    b.close()
    cw = info['counters'][t0 - 1:t1]
    out = {
        'metric': METRIC, 'value': value, 'unit': 'body-steps/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': 1e3 * wall / args.steps, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic',
        'config': {'workload': workload, 'worlds': args.worlds, 'bodies': n_bodies, 'dt': DT,
                   'solver_iterations': 10, 'sleep_enabled': True, 'seed': args.seed, 'window': [t0, t1],
                   'contacts_per_step_window': mean_counter(cw, 'contacts'),
                   'contacts_per_step_rollout': mean_counter(info['counters'], 'contacts')},
        'p50_step_ms': 1e3 * float(np.percentile(info['window_lat'], 50)),
        'p99_step_ms': 1e3 * float(np.percentile(info['window_lat'], 99)),
        'ms_per_step_synced': 1e3 * info['wall'] / args.steps, 'synced_equals_chained': same_synced,
        'contacts_per_s': mean_counter(cw, 'contacts') * args.steps / wall,
        'rollout': {'steps': total, 'ms_per_step': 1e3 * float(info['lat'].mean()),
                    'p50_step_ms': 1e3 * float(np.percentile(info['lat'], 50)),
                    'p99_step_ms': 1e3 * float(np.percentile(info['lat'], 99)),
                    'body_steps_per_s': n_bodies / float(info['lat'].mean())},
    }
    log('headline: %.3f ms/step over steps %d..%d, %.0f contacts/step (rollout mean %.0f)'
        % (out['ms_per_step'], t0, t1, out['config']['contacts_per_step_window'],
           out['config']['contacts_per_step_rollout']))
    if args.only_headline:
        print(json.dumps(out), flush=True)
        return

    if not args.no_timing:
        # the same window on a second batch, its steps launched eagerly with HIP
        # events between the stages on the batch stream (events cannot be timed
        # inside a captured graph)
        bt, ti = rollout(arrays, 'fp64', args.steps, args.warmup, timing=True)
        same = all(np.array_equal(p, q) for p, (q, _, _, _) in zip(snap64[:64], bt.snapshot(list(range(64)))))
        bt.close()
        kt, kb = ti['kernel_ms'], ti['kernel_bytes']
        K = args.steps
        kt['solve'] = kt['prepare'] + kt['solve_top'] + kt['solve_rest']
        out['eager_ms_per_step'] = 1e3 * ti['wall'] / K
        out['kernels_ms_per_step'] = {k: v / K for k, v in kt.items()}
        out['roofline'] = solve_roofline(ti, K, (t0, t1))
        step_ms = 1e3 * ti['wall'] / K
        sb = kb['step'] / K
        out['step_roofline'] = {'bytes_per_step': sb, 'ms_per_step': out['ms_per_step'],
                                'achieved_GBps': sb / (out['ms_per_step'] * 1e-3) / 1e9,
                                'frac': sb / (out['ms_per_step'] * 1e-3) / 1e9 / HBM_PEAK_GBS}
        out['stage_roofline'] = {k: {'avg_ms': kt[k] / K, 'GBps': kb[k] / max(kt[k] * 1e-3, 1e-12) / 1e9}
                                 for k in ('pre', 'sap', 'narrow', 'solve')}
        log('timed pass: %.2f ms/step eager (graph %.2f), identical results: %s' % (step_ms, out['ms_per_step'],
                                                                                   same))

    # the batched FP32 form on the same window
    b32, i32 = rollout(arrays, 'fp32', args.steps, args.warmup, record=False, chained=True)
    b32.close()
    b32, l32 = rollout(arrays, 'fp32', args.steps, args.warmup, record=False)
    b32.close()
    out['fp32_batched'] = {'value': n_bodies * args.steps / i32['wall'], 'ms_per_step': 1e3 * i32['wall'] / args.steps,
                           'p50_step_ms': 1e3 * float(np.percentile(l32['window_lat'], 50)),
                           'p99_step_ms': 1e3 * float(np.percentile(l32['window_lat'], 99)), 'dtype': 'f32',
                           'over_fp64': wall / i32['wall']}

    rng = np.random.default_rng(11)
    ps = np.sort(rng.choice(len(arrays), size=min(args.plain_sample, len(arrays)), replace=False))
    pv, plat = run_plain([arrays[i] for i in ps], args.steps, args.warmup, 'fp64')
    out['plain'] = {'value': pv, 'unit': 'body-steps/s', 'dtype': 'f64', 'worlds_sampled': len(ps),
                    'window': [t0, t1], 'p50_step_ms': 1e3 * float(np.percentile(plat, 50)),
                    'p99_step_ms': 1e3 * float(np.percentile(plat, 99)), 'batched_over_plain': value / pv}

    if not args.no_cpu:
        worlds, cwall = oracle_window(arrays, args.steps, args.warmup, CPU_THREADS)
        exact = sum(int(np.array_equal(ow.state()['position'], p)) for ow, p in zip(worlds, snap64))
        out['cpu_baseline'] = {'value': n_bodies * args.steps / cwall, 'unit': 'body-steps/s',
                               'cores': CPU_THREADS, 'kind': 'port',
                               'sample': 'every one of the %d worlds (%d bodies) over the same window (steps %d..%d, '
                                         'advanced untimed to it), one world per task on %d threads, %.2f s wall; '
                                         'an upper bound on the reference CPU path (C loops, no Python per-pair '
                                         'work)' % (len(arrays), n_bodies, t0, t1, CPU_THREADS, cwall)}
        del worlds
        cs = np.sort(rng.choice(len(arrays), size=min(args.curve_sample, len(arrays)), replace=False))
        out['err'] = {'fp64_worlds_bit_exact_at_window_end': exact, 'worlds': len(arrays),
                      'curves_worlds': len(cs),
                      'curves': error_curves([arrays[i] for i in cs], [1, 3, 10, 30, 100, 200, 300], CPU_THREADS)}
        log('parity: %d of %d worlds bit-exact at step %d' % (exact, len(arrays), t1))

    if not args.no_extras:
        out['serving'] = serving_run((400.0, 1600.0, 6400.0), 160, 'fp32')
        out['single_world'] = single_worlds(args.steps, args.warmup, not args.no_cpu)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
