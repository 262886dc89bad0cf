"""Benchmark: body-steps/s of the batched contact pipeline on one MI355X.

Workload (default): 4096 independent worlds, each a static ground box
(half extents 10 x 0.5 x 10) plus n_w dynamic bodies, n_w =
clip(round(lognormal(ln 64, 1.0)), 8, 2048) (seed 3, ~437k bodies), a 50/50
mix of unit boxes and r = 0.5 spheres of mass 1 dropped from y in [2, 50]
over x, z in [-10, 10]; mu = 0.5, e = 0, dt = 1/60, 10 solver iterations,
sleeping on.  One "step" advances every world by dt; the timed region is steps
1..K of the scene from its initial state (a separate batch takes the warm-up
steps), inputs resident in HBM, one graph launch per step.

Printed: one JSON line.  ``value`` is the batched FP64 form, which is
bit-identical to the FP64 oracle (the reference's algorithm) and so meets the
300-step position-error bound exactly; the batched FP32 form (FP32
narrowphase and solver) is reported beside it with its errors against FP64
(after one step from the same state, and after 300 steps, next to the
divergence that a 1-ulp change of the initial positions causes in FP64 alone
-- the piles are chaotic).  Also: the plain one-world-at-a-time GPU form,
p50/p99 step latency, per-stage device time, the roofline of the dominant
kernel, and the CPU baseline (the oracle, a C restatement of the reference
algorithm, one world per thread on this host's cores).
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

METRIC = 'body-steps/s at 1 GPU (batched FP32 vs plain) + p99 step latency, err≤1e-4m'
DT = 1.0 / 60.0
HBM_PEAK_GBS = 8000.0


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


def algorithmic_bytes(kernel, c, r):
    """HBM bytes one launch of ``kernel`` must move, from the step's counters.

    Per-unit figures (r = bytes of the narrowphase/solver float):
      pre     350 + 18r per body (pose, velocity and parameters in; previous pose,
              velocity, AABB, rotation and world inverse inertia out)
      sap     16 per body (lo.x, hi.x) + 32 per x-window candidate (y/z bounds)
              + 8 per solid pair written
      narrow  106 + 18r per solid pair (indices, kinds, poses, rotations,
              half extents) + 24 + r per contact point written
      solve   32 + 7r per contact (indices, point, normal, depth, impulses)
              + 152 + 9r per body touched (velocity, pose, mass and inertia in
              and out), bodies touched taken as min(bodies, 2 x manifolds)
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
    """Algorithmic bytes of one solver launch of the largest size class present.

    Per contact: its solver row (224 B, both precisions) in and its three
    impulses (3r) out; per manifold its 16-byte record; per body in contact
    (taken as min(bodies, 2 x manifolds) per world): FP64 velocity in and out
    (96 B), FP64 position in and out (48 B) and world inverse inertia (12r).
    """
    top = int(stats['size_class'].max())
    sel = stats['size_class'] == top
    k = int(stats['contacts'][sel].sum())
    m = int(stats['manifolds'][sel].sum())
    nd = int(np.minimum(stats['bodies'][sel], 2 * stats['manifolds'][sel]).sum())
    return k * (224 + 3 * r) + 16 * m + nd * (96 + 48 + 12 * r), top, int(sel.sum())


def run_batched(arrays, precision, steps, warmup, timing=True, lanes=False):
    """Warm up on one batch, then time ``steps`` steps of a fresh batch from the initial state.

    With ``timing`` the steps launch eagerly with a HIP event between kernels on
    each lane's stream (events cannot be timed inside a captured graph); without
    it every step is one graph launch.  Stage times, counters and algorithmic
    bytes come back per lane (one lane for a single batch).
    """
    from omi_physics_amd.batch import batch_from_arrays
    if warmup > 0:
        w = batch_from_arrays(arrays, precision=precision, lanes=lanes)
        w.prepare(DT)
        w.step(DT, warmup)
        w.sync()
        w.close()
    t0 = time.perf_counter()
    b = batch_from_arrays(arrays, precision=precision, lanes=lanes)
    log('%s batch loaded: %.1f GB device, %.1f s' % (precision, b.device_bytes() / 1e9, time.perf_counter() - t0))
    if timing:
        b.set_timing(True)
    b.prepare(DT)
    b.sync()
    lat = []
    lanes = b.lanes if hasattr(b, 'lanes') else [b]
    ktot = [{} for _ in lanes]
    ctot = [{} for _ in lanes]
    per_kernel_bytes = [{} for _ in lanes]
    r = 4 if precision == 'fp32' else 8
    t_start = time.perf_counter()
    for s in range(steps):
        t = time.perf_counter()
        b.step(DT, 1)
        b.sync()
        lat.append(time.perf_counter() - t)
        if timing:
            for li, lb in enumerate(lanes):
                kt = lb.kernel_times()
                c = lb.counters()
                for k, v in kt.items():
                    ktot[li][k] = ktot[li].get(k, 0.0) + v
                for k, v in c.items():
                    ctot[li][k] = ctot[li].get(k, 0) + v
                kb = per_kernel_bytes[li]
                for k in ('pre', 'sap', 'narrow', 'solve'):
                    kb[k] = kb.get(k, 0) + algorithmic_bytes(k, c, r)
                tb, top, ntop = solve_top_bytes(lb.world_stats(), r)
                kb['solve_top'] = kb.get('solve_top', 0) + tb
                kb['top_class'] = (top, ntop)
        if (s + 1) % 50 == 0:
            log('  %s step %d/%d  %.2f ms' % (precision, s + 1, steps, 1e3 * lat[-1]))
    wall = time.perf_counter() - t_start
    return b, np.array(lat), wall, ktot, ctot, per_kernel_bytes


def run_plain(arrays, steps, precision):
    """The plain form: each sampled world alone on the GPU, one world at a time."""
    from omi_physics_amd.batch import batch_from_arrays
    body_steps = 0
    total = 0.0
    lat = []
    for a in arrays:
        b = batch_from_arrays([a], precision=precision)
        b.prepare(DT)
        b.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            t = time.perf_counter()
            b.step(DT, 1)
            b.sync()
            lat.append(time.perf_counter() - t)
        total += time.perf_counter() - t0
        body_steps += a['n'] * steps
        b.close()
    return body_steps / total, np.array(lat)


def cpu_baseline(arrays, steps, threads):
    """The oracle (C, FP64, one world per thread) on a sample of the worlds."""
    import oracle

    def one(a):
        ow = oracle.OracleWorld(a)
        ow.step(DT, steps)
        return ow.state()['position']

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        pos = list(ex.map(one, arrays))
    wall = time.perf_counter() - t0
    return sum(a['n'] for a in arrays) * steps / wall, wall, pos


def single_world(name, arrays, precision, steps, warmup, cpu):
    """One world stepped alone on the GPU (the plain form), vs the oracle on one core."""
    from omi_physics_amd.batch import batch_from_arrays
    n = int(arrays['n'])
    if warmup:
        w = batch_from_arrays([arrays], precision=precision)
        w.step(DT, warmup); w.sync(); w.close()
    b = batch_from_arrays([arrays], precision=precision)
    b.prepare(DT); b.sync()
    lat = []
    cmax = 0
    t0 = time.perf_counter()
    for s in range(steps):
        t = time.perf_counter()
        b.step(DT, 1); b.sync()
        lat.append(time.perf_counter() - t)
        cmax = max(cmax, b.counters()['contacts'])
    wall = time.perf_counter() - t0
    pos = b.read(0)['position']
    b.close()
    out = {'metric': METRIC, 'value': n * steps / wall, 'unit': 'body-steps/s', 'n_gpus': 1, 'steps': steps,
           'warmup': warmup, 'ms_per_step': 1e3 * wall / steps, 'higher_is_better': True, 'scaling': 'weak',
           'vs_baseline': None, 'dtype': 'f64' if precision == 'fp64' else 'f32', 'data': 'synthetic',
           'config': {'workload': name, 'bodies': n, 'dt': DT},
           'p50_step_ms': 1e3 * float(np.percentile(lat, 50)), 'p99_step_ms': 1e3 * float(np.percentile(lat, 99)),
           'max_contacts': cmax}
    if cpu:
        import oracle
        ow = oracle.OracleWorld(arrays)
        t0 = time.perf_counter()
        ow.step(DT, steps)
        cw = time.perf_counter() - t0
        err = float(np.abs(ow.state()['position'] - pos).max())
        out['cpu_baseline'] = {'value': n * steps / cw, 'unit': 'body-steps/s', 'cores': 1, 'kind': 'port',
                               'sample': 'the whole world, %d steps, one core (%.1f s)' % (steps, cw)}
        out['err'] = {'max_pos_m_vs_fp64_oracle': err}
    return out


def one_step_error(worlds, states):
    """FP32 against FP64 after one step from the same state (the worlds at step K):
    the precision of the FP32 narrowphase + solver before any chaotic growth."""
    import oracle
    from omi_physics_amd.batch import batch_from_arrays
    start = []
    for a, st in zip(worlds, states):
        a = dict(a)
        for k_arr, k_st in (('position', 'position'), ('orientation', 'orientation'),
                            ('linear_velocity', 'linear_velocity'), ('angular_velocity', 'angular_velocity'),
                            ('sleep_timer', 'sleep_timer')):
            a[k_arr] = np.ascontiguousarray(st[k_st], dtype=np.float64)
        a['awake'] = np.ascontiguousarray(st['awake'], dtype=np.uint8)
        start.append(a)
    b = batch_from_arrays(start, precision='fp32')
    b.step(DT, 1)
    b.sync()
    ep, ev = 0.0, 0.0
    for i, a in enumerate(start):
        ow = oracle.OracleWorld(a)
        ow.step(DT, 1)
        o, g = ow.state(), b.read(i)
        ep = max(ep, float(np.abs(o['position'] - g['position']).max()))
        ev = max(ev, float(np.abs(o['linear_velocity'] - g['linear_velocity']).max()))
    b.close()
    return {'fp32_one_step_pos_max_m': ep, 'fp32_one_step_vel_max_mps': ev}


def serving(rates, n_requests, precision, seed=5):
    from omi_physics_amd import scenes
    from omi_physics_amd.serve import Request, RolloutServer, run_plain
    res = {}
    for rate in rates:
        reqs_spec = scenes.serving_requests(n_requests, rate, seed=seed)
        reqs = [Request(id=i, arrays=scenes.arrays(scenes.boxes_on_ground(r['n'], seed=r['seed'],
                                                                              sphere_fraction=0.5)),
                        steps=r['steps'], arrival=r['arrival']) for i, r in enumerate(reqs_spec)]
        srv = RolloutServer(max_worlds=1024, max_bodies=1 << 19, precision=precision)
        st = srv.run(reqs).summary()
        srv.close()
        wait = [r.admitted_at - r.arrival for r in reqs]
        st['queue_wait_p99_s'] = float(np.percentile(wait, 99))
        res[str(rate)] = st
        log('serving rate %s/s: %s' % (rate, st))
    plain_reqs = [Request(id=i, arrays=scenes.arrays(scenes.boxes_on_ground(r['n'], seed=r['seed'],
                                                                            sphere_fraction=0.5)),
                          steps=r['steps']) for i, r in enumerate(scenes.serving_requests(24, 1.0, seed=seed))]
    res['plain'] = run_plain(plain_reqs, precision=precision)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--workload', default='batched4096',
                    choices=['batched4096', 'boxes1000', 'pyramids', 'convex', 'serving', 'copy'])
    ap.add_argument('--copy-mib', type=int, default=1024, help='buffer size of the copy workload')
    ap.add_argument('--precision', default=None, help='fp32 or fp64 for the single-world workloads')
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=300)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--worlds', type=int, default=4096)
    ap.add_argument('--seed', type=int, default=3)
    ap.add_argument('--plain-sample', type=int, default=48, help='worlds stepped alone for the plain form')
    ap.add_argument('--cpu-sample', type=int, default=192, help='worlds the CPU baseline and error check step')
    ap.add_argument('--no-plain', action='store_true')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-fp32', action='store_true', help='skip the batched FP32 form')
    ap.add_argument('--no-timing', action='store_true')
    ap.add_argument('--lanes', action='store_true', help='split the worlds by size class into lanes stepped side by side')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU only: there is no multi-GPU path (independent worlds would shard '
              'across GPUs by world id, with no collective on the step path).', file=sys.stderr)
        sys.exit(2)
    from omi_physics_amd import _ffi
    if _ffi.load().omi_device_count() < 1:
        print('no HIP device visible', file=sys.stderr)
        sys.exit(1)
    if args.workload == 'copy':
        # a plain device-to-device copy of known size through the HIP runtime (no
        # framework): the achievable HBM rate and the calibration of the
        # FETCH_SIZE / WRITE_SIZE counters
        import ctypes
        hip = ctypes.CDLL(os.path.join(os.environ.get('ROCM_PATH', '/opt/rocm'), 'lib', 'libamdhip64.so'))
        nbytes = ctypes.c_size_t(args.copy_mib << 20)
        x, y = ctypes.c_void_p(), ctypes.c_void_p()
        if hip.hipMalloc(ctypes.byref(x), nbytes) or hip.hipMalloc(ctypes.byref(y), nbytes):
            raise RuntimeError('hipMalloc failed')
        hip.hipMemset(x, 1, nbytes)
        d2d = 3  # hipMemcpyDeviceToDevice
        for _ in range(3):
            hip.hipMemcpyAsync(y, x, nbytes, d2d, None)
        hip.hipDeviceSynchronize()
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        hip.hipEventCreate(ctypes.byref(e0))
        hip.hipEventCreate(ctypes.byref(e1))
        reps = 20
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(y, x, nbytes, d2d, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        per = ms.value / reps
        print(json.dumps({'workload': 'copy', 'bytes_read_per_copy': nbytes.value, 'bytes_written_per_copy':
                          nbytes.value, 'ms_per_copy': per, 'GBps': 2 * nbytes.value / (per * 1e-3) / 1e9}),
              flush=True)
        hip.hipFree(x)
        hip.hipFree(y)
        return
    if args.workload != 'batched4096':
        from omi_physics_amd import scenes, PhysicsWorld, model
        if args.workload == 'serving':
            out = {'metric': METRIC, 'workload': 'serving', 'unit': 'body-steps/s',
                   'rates': serving([100.0, 400.0, 1600.0], 160, args.precision or 'fp32')}
            print(json.dumps(out), flush=True)
            return
        if args.workload == 'boxes1000':
            arr = scenes.arrays(scenes.boxes_on_ground(1000, seed=1))
            prec = args.precision or 'fp64'
            name = '1000 unit boxes on a 20x20 m ground, one world, %s' % prec
        elif args.workload == 'pyramids':
            arr = scenes.arrays(scenes.pyramids_and_debris())
            prec = args.precision or 'fp32'
            name = '20 ten-level box pyramids + 8000 box/sphere debris, one world, %s' % prec
        else:
            arr = scenes.build(scenes.convex_and_chains(), PhysicsWorld, model).export_arrays()
            prec = args.precision or 'fp32'
            name = ('2000 cooked convex hulls (16-32 points) and 20 hinge chains of 50 links on a ground box, '
                    'one world, %s' % prec)
        out = single_world(name, arr, prec, args.steps, args.warmup, not args.no_cpu)
        print(json.dumps(out), flush=True)
        return

    arrays = make_worlds(args.worlds, args.seed)
    n_bodies = int(sum(a['n'] for a in arrays))
    rng = np.random.default_rng(11)
    sample = np.sort(rng.choice(len(arrays), size=min(args.cpu_sample, len(arrays)), replace=False))

    # headline: the batched FP64 form, graph-launched steps 1..K
    b, lat, wall, _, _, _ = run_batched(arrays, 'fp64', args.steps, args.warmup, timing=False, lanes=args.lanes)
    value = n_bodies * args.steps / wall
    gpu64 = {int(i): b.read(int(i)) for i in sample}
    b.close()
    out = {
        'metric': METRIC, 'value': value, 'unit': 'body-steps/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': 1e3 * wall / args.steps, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic',
        'config': {'workload': '%d independent box/sphere worlds, lognormal(ln 64, 1.0) bodies clipped [8, 2048], '
                               'batched in one graph-captured step%s, FP64 (bit-identical to the oracle); steps '
                               '1..%d from rest' % (args.worlds, ' per size-class lane' if args.lanes else '', args.steps),
                   'worlds': args.worlds, 'bodies': n_bodies, 'dt': DT, 'solver_iterations': 10,
                   'sleep_enabled': True, 'seed': args.seed, 'lanes': bool(args.lanes)},
        'p50_step_ms': 1e3 * float(np.percentile(lat, 50)),
        'p99_step_ms': 1e3 * float(np.percentile(lat, 99)),
    }
    if not args.no_timing:
        # the same K steps again from the initial state, launched eagerly with HIP
        # events between the stages on the batch stream: device time per stage,
        # counters and the roofline of the dominant kernel
        bt, lat_t, wall_t, ktot, ctot, kbytes = run_batched(arrays, 'fp64', args.steps, 0, timing=True,
                                                            lanes=args.lanes)
        same = all(np.array_equal(bt.read(int(i))['position'], gpu64[int(i)]['position']) for i in sample[:8])
        bt.close()
        log('timed pass: %.2f ms/step eager (graph %.2f), identical results: %s'
            % (1e3 * wall_t / args.steps, 1e3 * wall / args.steps, same))
        out['eager_ms_per_step'] = 1e3 * wall_t / args.steps
        tot = {}
        for c in ctot:
            for k, v in c.items():
                if k not in ('overflow', 'max_levels'):
                    tot[k] = tot.get(k, 0) + int(v)
        out['contacts_per_s'] = tot.get('contacts', 0) / wall
        out['totals'] = tot
        # one lane per size class present (the worlds of each class are a member
        # batch of one group; the lanes' steps run side by side)
        names = []
        for li, kt in enumerate(ktot):
            kt['solve'] = kt['prepare'] + kt['solve_top'] + kt['solve_rest']
            cls, nw = kbytes[li]['top_class']
            names.append('lane %d: %d worlds of size class %d' % (li, nw, cls))
        out['kernels_ms_per_step'] = {nm: {k: v / args.steps for k, v in kt.items()} for nm, kt in zip(names, ktot)}
        crit = int(np.argmax([kt['solve_top'] for kt in ktot]))
        top, ntop = kbytes[crit]['top_class']
        avg_ms = ktot[crit]['solve_top'] / args.steps
        per_launch = kbytes[crit]['solve_top'] / args.steps
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        traffic = None
        tf = os.path.join(HERE, 'profiles', 'traffic.json')
        if os.path.exists(tf):
            with open(tf) as f:
                traffic = json.load(f).get('k_solve_top_fp64_bytes_per_launch')
        out['roofline'] = {'kernel': 'k_solve<double> size class %d (%d worlds, the largest; alone on its '
                                     'stream)' % (top, ntop),
                           'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                           'frac': achieved / HBM_PEAK_GBS, 'traffic': traffic, 'avg_ms': avg_ms,
                           'bytes_per_launch': per_launch}
        out['roofline_stages'] = {nm: {k: {'avg_ms': kt[k] / args.steps,
                                           'GBps': kb[k] / max(kt[k] * 1e-3, 1e-12) / 1e9}
                                       for k in ('pre', 'sap', 'narrow', 'solve')}
                                  for nm, kt, kb in zip(names, ktot, kbytes)}

    if not args.no_fp32:
        b32, lat32, wall32, *_ = run_batched(arrays, 'fp32', args.steps, 0, timing=False, lanes=args.lanes)
        gpu32 = {int(i): b32.read(int(i))['position'] for i in sample}
        b32.close()
        out['fp32_batched'] = {'value': n_bodies * args.steps / wall32, 'ms_per_step': 1e3 * wall32 / args.steps,
                               'p50_step_ms': 1e3 * float(np.percentile(lat32, 50)),
                               'p99_step_ms': 1e3 * float(np.percentile(lat32, 99)), 'dtype': 'f32',
                               'over_fp64': wall / wall32}

    if not args.no_plain:
        ps = np.sort(rng.choice(len(arrays), size=min(args.plain_sample, len(arrays)), replace=False))
        pv, plat = run_plain([arrays[i] for i in ps], args.steps, 'fp64')
        out['plain'] = {'value': pv, 'unit': 'body-steps/s', 'dtype': 'f64', 'worlds_sampled': len(ps),
                        'p50_step_ms': 1e3 * float(np.percentile(plat, 50)),
                        'p99_step_ms': 1e3 * float(np.percentile(plat, 99)),
                        'batched_over_plain': value / pv}

    if not args.no_cpu:
        threads = min(len(os.sched_getaffinity(0)), 16)
        sub = [arrays[i] for i in sample]
        cv, cwall, cpos = cpu_baseline(sub, args.steps, threads)
        out['cpu_baseline'] = {'value': cv, 'unit': 'body-steps/s', 'cores': threads, 'kind': 'port',
                               'sample': '%d of the %d worlds (%d bodies), %d steps each, one world per '
                                         'thread, %.1f s wall' % (len(sub), len(arrays),
                                                                   sum(a['n'] for a in sub), args.steps, cwall)}
        e64 = np.array([np.abs(gpu64[int(i)]['position'] - p).max() for i, p in zip(sample, cpos)])
        err = {'worlds': len(sample), 'steps': args.steps, 'fp64_max_m': float(e64.max())}
        if not args.no_fp32:
            # FP64 alone: the same worlds with every dynamic body's initial position
            # moved by one ulp -- how far the chaotic piles carry a last-bit change
            def nudged(a):
                a = dict(a)
                p = a['position'].copy()
                dyn = a['motion_type'] == 2
                p[dyn] = np.nextafter(p[dyn], np.inf)
                a['position'] = p
                return a
            _, _, upos = cpu_baseline([nudged(a) for a in sub], args.steps, threads)
            e32 = np.array([np.abs(gpu32[int(i)] - p).max() for i, p in zip(sample, cpos)])
            eu = np.array([np.abs(u - p).max() for u, p in zip(upos, cpos)])
            err.update({'fp32_max_m': float(e32.max()), 'fp32_median_world_max_m': float(np.median(e32)),
                        'fp32_worlds_within_1e-4': float(np.mean(e32 <= 1e-4)),
                        'fp64_1ulp_max_m': float(eu.max()), 'fp64_1ulp_median_world_max_m': float(np.median(eu)),
                        'fp64_1ulp_worlds_within_1e-4': float(np.mean(eu <= 1e-4))})
            err.update(one_step_error(sub, [gpu64[int(i)] for i in sample]))
        out['err'] = err
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
