#!/usr/bin/env python
"""Benchmark of the dense-tracking hot path on one MI355X.

Workload ``global300``: 300 keyframes at 48x64, 3000 edges (Poisson(10) out-degree),
one ``ba`` call of 8 Gauss-Newton iterations (t0 = 0, t1 = 300, lm = 1e-4, ep = 0.1)
plus one bf16 4-level radius-3 CorrBlock lookup over all 3000 edges per iteration.
A step = that ba call + 8 lookups; ``value`` = GN iterations per second.

Beside it, in the same process: BA-only iterations/s, lookup GB/s, the plain
all-FP32 path (FP32 reduced Cholesky and back-substitution, FP32 volume) and the
final weighted RMS reprojection error of both paths from the same start; the
oracle timed on the host as the CPU baseline.  Inputs are synthetic and already
resident in HBM when timing starts.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'hi-slam2_amd'))

HBM_PEAK_GBPS = 8000.0
LOOKUP_BYTES_BF16 = 912      # 4 levels * 8*8 taps * 2 B + 196 outputs * 2 B + 8 B coords
LOOKUP_BYTES_F32 = 1816
METRIC = 'dense-BA GN iters/s & corr-lookup GB/s vs plain GPU path at equal reproj error'


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-plain', action='store_true', help='skip the all-FP32 comparison path')
    ap.add_argument('--copy-probe', action='store_true',
                    help='also run one 2 GiB device copy (counter calibration for rocprofv3 --pmc)')
    ap.add_argument('--only', choices=['all', 'ba', 'lookup', 'altcorr', 'window', 'hires64', 'mono', 'geometry',
                                       'pgba'],
                    default='all',
                    help='ba: time only the global300 ba call (split precision; short runs for rocprofv3 counters); '
                         'lookup: time only the bf16 lookup (short runs for rocprofv3 counters); '
                         'mono: time the mono-prior step (JDSA, proj_trans) on the global300 graph; '
                         'geometry: time frame_distance, depth_filter and iproj on it; '
                         'pgba: time the Sim3 pose-graph BA step on a 300-keyframe loop-closure graph; '
                         'altcorr: time the on-the-fly correlation (AltCorrBlock) on global300 and hires64; '
                         'window: latency of one frontend-window ba call (12 keyframes, 70 edges, 4 iterations); '
                         'hires64: 96x128 BA (64 keyframes, 800 edges, 4 iterations) + radius-4 on-the-fly lookup')
    return ap.parse_args()


def log(*a):
    print('[bench]', *a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f'bench.py measures one GPU only; --gpus {args.gpus} is not supported.', file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from hislam2_amd import backend, synth
    from hislam2_amd import corr as hcorr
    from hislam2_amd.metrics import reprojection_rms
    db = backend()
    dev = torch.device('cuda:0')
    torch.cuda.set_device(dev)

    if args.only == 'pgba':
        pgba_bench(db, dev, args)
        return
    if args.only == 'altcorr':
        altcorr_bench(dev, args)
        return
    if args.only == 'window':
        window_bench(db, dev, args)
        return
    if args.only == 'hires64':
        hires64_bench(db, dev, args)
        return
    log('generating global300')
    p = synth.ba_global300()
    E, iters = len(p.ii), p.iterations
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    poses0, disps0 = up(p.poses), up(p.disps)
    intr, tg, wt, eta, ii, jj = up(p.intrinsics), up(p.targets), up(p.weights), up(p.eta), up(p.ii), up(p.jj)
    fm32 = up(synth.fmaps(7, p.poses.shape[0]))             # [1, 300, 128, 48, 64]
    coords = up(synth.lookup_coords(8, E))                    # [1, 3000, 48, 64, 2]
    h, w = p.disps.shape[1:]
    lookups = E * h * w

    if args.copy_probe:
        a = torch.empty(1 << 29, dtype=torch.float32, device=dev)   # 2 GiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    if args.only == 'ba':
        poses, disps = poses0.clone(), disps0.clone()

        def call():
            return db.ba(poses, disps, intr, tg, wt, eta, ii, jj, p.t0, p.t1, iters, p.lm, p.ep, False)

        ms = time_launches(call, args.steps, args.warmup)
        print(json.dumps({'metric': 'BA-only GN it/s (global300, split precision)',
                          'value': round(iters / (ms / 1e3), 3), 'unit': 'GN it/s', 'ms_per_call': round(ms, 4),
                          'iterations_per_call': iters, 'edges': E}), flush=True)
        return

    if args.only == 'lookup':
        f = fm32.to(torch.bfloat16)
        blk = hcorr.CorrBlock(f[:, p.ii], f[:, p.jj], dtype=torch.bfloat16)
        del f
        for _ in range(args.warmup):
            blk(coords)
        torch.cuda.synchronize()
        ev = []
        for _ in range(args.steps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            blk(coords)
            e1.record()
            ev.append((e0, e1))
        torch.cuda.synchronize()
        ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
        gbps = LOOKUP_BYTES_BF16 * lookups / (ms / 1e3) / 1e9
        # the lookup with corr_encoder's first layer (1x1 conv 196 -> 128 + ReLU): fused, and as
        # the lookup followed by torch's bf16 convolution
        import torch.nn.functional as F
        g = torch.Generator().manual_seed(0)
        W = (torch.randn(128, 196, 1, 1, generator=g) / 14.0).to(dev).to(torch.bfloat16)
        bias = torch.zeros(128, device=dev)
        B_, N_, h_, w_ = coords.shape[:4]

        def fused():
            return blk.lookup_conv1x1(coords, W, bias)

        def unfused():
            return F.relu(F.conv2d(blk(coords).view(B_ * N_, 196, h_, w_), W, bias.to(torch.bfloat16)))

        conv_ms = {}
        for name, fn in (('lookup_conv_fused_ms', fused), ('lookup_then_conv_ms', unfused)):
            for _ in range(max(args.warmup, 1)):
                fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(args.steps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            conv_ms[name] = round(e0.elapsed_time(e1) / args.steps, 4)
        print(json.dumps({'metric': 'corr-lookup GB/s (algorithmic, 912 B per bf16 lookup)',
                          'value': round(gbps, 1), 'unit': 'GB/s', 'lookup_ms': round(ms, 4),
                          'lookups_per_launch': lookups, 'tiled': blk.tiled, **conv_ms}), flush=True)
        return

    if args.only == 'mono':
        mono_bench(db, p, up, dev, args)
        return
    if args.only == 'geometry':
        geometry_bench(db, p, up, args)
        return

    def ba_call(poses, disps, precision):
        return db.ba_ex(poses, disps, intr, tg, wt, eta, ii, jj, p.t0, p.t1, iters, p.lm, p.ep,
                        False, precision)

    def ba_timed(poses, disps, precision):
        # the reference's entry point for the split path (no status read-back, so the host
        # queues the lookups behind it); ba_ex for the all-FP32 comparison path
        if precision == 0:
            return db.ba(poses, disps, intr, tg, wt, eta, ii, jj, p.t0, p.t1, iters, p.lm, p.ep, False)
        return ba_call(poses, disps, precision)

    # ---- accuracy: both paths from the same start, against the float64 oracle ----------
    res = {}
    for name, prec in (('split', 0), ('plain', 1)):
        if name == 'plain' and args.no_plain:
            continue
        P_, D_ = poses0.clone(), disps0.clone()
        out = ba_call(P_, D_, prec)
        torch.cuda.synchronize()
        res[name] = dict(ok=int(out[3].item()),
                         rms=reprojection_rms(P_, D_, intr, tg, wt, ii, jj),
                         oracle=ba_oracle_errors(P_, D_, out[0], out[1]))
    rms0 = reprojection_rms(poses0, disps0, intr, tg, wt, ii, jj)
    log('reprojection rms start', rms0, res)

    def time_path(dtype, precision, check_lookup=False):
        f = fm32.to(dtype)
        f1, f2 = f[:, p.ii], f[:, p.jj]
        del f
        torch.cuda.synchronize()
        # volume build, timed on its own: the first one includes the allocation of the pyramid
        # (cold), the rebuild reuses the freed memory through torch's caching allocator, as a
        # tracking system rebuilding the volume of each global BA does (steady state)
        build = []
        for _ in range(2):
            b0, b1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            blk = None
            b0.record()
            blk = hcorr.CorrBlock(f1, f2, dtype=dtype)
            b1.record()
            torch.cuda.synchronize()
            build.append(b0.elapsed_time(b1))
        build_cold_ms, build_ms = build
        del f1, f2
        poses, disps = poses0.clone(), disps0.clone()
        stream = torch.cuda.current_stream()
        ev = []

        def step(record):
            ba_timed(poses, disps, precision)
            for _ in range(iters):
                if record:
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    blk(coords)
                    e1.record(stream)
                    ev.append((e0, e1))
                else:
                    blk(coords)

        for _ in range(args.warmup):
            step(False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(True)
        torch.cuda.synchronize()
        t_total = time.perf_counter() - t0
        lookup_ms = [a.elapsed_time(b) for a, b in ev]
        lookup_err = lookup_oracle_errors(blk(coords)) if check_lookup else None
        # BA alone
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            ba_timed(poses, disps, precision)
        torch.cuda.synchronize()
        t_ba = time.perf_counter() - t1
        del blk
        torch.cuda.empty_cache()
        return dict(t_total=t_total, t_ba=t_ba, lookup_ms=float(np.mean(lookup_ms)), build_ms=build_ms,
                    build_cold_ms=build_cold_ms,
                    lookup_err=lookup_err)

    log('timing split path (bf16 volume, FP64 solve)')
    split = time_path(torch.bfloat16, 0, check_lookup=True)
    plain = None
    if not args.no_plain:
        log('timing plain path (fp32 volume, FP32 solve)')
        plain = time_path(torch.float32, 1)

    K = args.steps
    value = iters * K / split['t_total']
    # corr build, per edge: 2*HW*C*2 B of bf16 feature maps + HW^2 * (1 + 1/4 + 1/16 + 1/64) * 2 B
    HWp = h * w
    build_bytes = E * (2 * HWp * fm32.shape[2] * 2 + HWp * HWp * (1 + 0.25 + 0.0625 + 0.015625) * 2)
    lookup_s = split['lookup_ms'] / 1e3
    achieved = LOOKUP_BYTES_BF16 * lookups / lookup_s / 1e9
    traffic = lookup_traffic()

    # the on-the-fly correlation (AltCorrBlock, the reference's global-BA path) in both
    # precisions: the same MFMA kernel with bf16 and with fp32 operands
    onthefly = None
    if not args.no_plain:
        onthefly = {}
        for dname, dt, key in (('bf16', torch.bfloat16, 'split'), ('fp32', torch.float32, 'plain')):
            ablk = hcorr.AltCorrBlock(fm32, dtype=dt)
            ms = time_launches(lambda: ablk(coords, ii, jj), 5, 1)
            t_ba = (split if key == 'split' else plain)['t_ba'] / args.steps
            onthefly[dname] = {'lookup_ms': round(ms, 4),
                               'GN_it_s': round(iters / (t_ba + iters * ms / 1e3), 3)}
            del ablk
            torch.cuda.empty_cache()
        onthefly['speedup_vs_plain'] = round(onthefly['bf16']['GN_it_s'] / onthefly['fp32']['GN_it_s'], 3)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p, fm32, coords)

    line = {
        'metric': METRIC,
        'value': round(value, 3),
        'unit': 'GN it/s',
        'n_gpus': 1,
        'steps': K,
        'warmup': args.warmup,
        'ms_per_step': round(1e3 * split['t_total'] / K, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'bf16 corr / f32 jacobians+hessians / f64 reduced solve',
        'data': 'synthetic (seeded random-walk poses, U(0.2,1) inverse depth, N(0,1)/sqrt(128) fmaps)',
        'config': {'workload': 'global300: 300 keyframes 48x64, 3000 edges, ba(8 GN iterations, '
                               't0=0, t1=300, lm=1e-4, ep=0.1) + bf16 4-level r=3 CorrBlock lookup '
                               'of all 3000 edges per iteration',
                   'edges': E, 'keyframes': int(p.poses.shape[0]), 'iterations_per_call': iters},
        'ba_only_it_s': round(iters * K / split['t_ba'], 3),
        'lookup_ms': round(split['lookup_ms'], 4),
        'corr_build_ms': round(split['build_ms'], 3),
        'corr_build_cold_ms': round(split['build_cold_ms'], 3),
        'corr_build_GBps_algorithmic': round(build_bytes / (split['build_ms'] / 1e3) / 1e9, 1),
        'lookup_GBps_algorithmic': round(achieved, 1),
        'reproj_rms_start_px': round(rms0, 6),
        'reproj_rms_split_px': round(res['split']['rms'], 6),
        'oracle_err_split': res['split']['oracle'],
        'oracle_err_lookup_bf16': split['lookup_err'],
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBPS,
                     'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBPS, 4), 'traffic': traffic,
                     'kernel': 'corr_lookup_vec4_kernel<bf16> (all 4 levels per block, r=3, 9.216M '
                               'lookups/launch, 912 B/lookup algorithmic)'},
        'cpu_baseline': cpu,
    }
    if plain is not None:
        pv = iters * K / plain['t_total']
        line.update({
            'plain_fp32_it_s': round(pv, 3),
            'plain_fp32_ba_only_it_s': round(iters * K / plain['t_ba'], 3),
            'plain_fp32_lookup_ms': round(plain['lookup_ms'], 4),
            'speedup_vs_plain': round(value / pv, 3),
            'reproj_rms_plain_px': round(res['plain']['rms'], 6),
            'reproj_diff_px': round(abs(res['split']['rms'] - res['plain']['rms']), 7),
            'oracle_err_plain': res['plain']['oracle'],
            'target_2x_met': bool(value / pv >= 2.0 and abs(res['split']['rms'] - res['plain']['rms']) <= 1e-3),
            'plain_path': 'the same kernels instantiated in FP32 (per-edge Hessian blocks, frame sums, Schur '
                          'Gram totals, reduced Cholesky, back-substitution) and the fp32 stored-volume lookup',
            'onthefly_corr': onthefly,
        })
    print(json.dumps(line), flush=True)


MFMA_PEAK_TFLOPS = {'bfloat16': 2500.0, 'float16': 2500.0, 'float32': 157.3}


def time_launches(fn, steps, warmup):
    """Mean milliseconds of fn() over `steps` launches, HIP events on torch's current stream (the
    stream the extension launches on)."""
    import torch
    for _ in range(max(warmup, 1)):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def altcorr_bench(dev, args):
    """The on-the-fly correlation (AltCorrBlock: feature pyramid built once, all levels per call in
    one launch on the matrix cores) over every edge of two workloads:
      global300 -- 300 keyframes at 48x64, 3000 edges, radius 3, 4 levels (196 channels);
      hires64   -- 64 keyframes at 96x128, 800 edges, radius 4, 4 levels (324 channels);
    coords = pixel grid + N(0, 2 px) flow.  Per lookup (one pixel-edge): algorithmic flops
    L (2r+2)^2 2C; algorithmic bytes = the feature maps once per pass, F HW C s (1 + 1.328), plus
    the outputs L (2r+1)^2 s_out.  One JSON line per workload and dtype."""
    import numpy as np
    import torch
    from hislam2_amd import corr as hcorr
    from hislam2_amd import synth
    work = [('global300', synth.ba_global300, 3, 7, 8), ('hires64', synth.ba_hires64, 4, 64, 65)]
    for name, gen, r, fseed, cseed in work:
        p = gen()
        n, h, w = p.disps.shape
        E = len(p.ii)
        fm = torch.from_numpy(synth.fmaps(fseed, n, h=h, w=w)).to(dev)
        coords = torch.from_numpy(synth.lookup_coords(cseed, E, h=h, w=w)).to(dev)
        ii, jj = torch.from_numpy(p.ii).to(dev), torch.from_numpy(p.jj).to(dev)
        lookups = E * h * w
        L, C, D, R = 4, fm.shape[2], 2 * r + 2, 2 * r + 1
        dtypes = [torch.bfloat16, torch.float32] + ([torch.float8_e4m3fn] if name == 'hires64' else [])
        for dt in dtypes:
            build_ms = time_launches(lambda: hcorr.AltCorrBlock(fm, radius=r, dtype=dt), 3, 1)
            blk = hcorr.AltCorrBlock(fm, radius=r, dtype=dt)
            ms = time_launches(lambda: blk(coords, ii, jj), args.steps, args.warmup)
            s_in = 1 if dt == torch.float8_e4m3fn else torch.tensor([], dtype=dt).element_size()
            s_out = 2 if dt == torch.float8_e4m3fn else s_in
            nbytes = n * h * w * C * s_in * (1 + 1.328) + lookups * L * R * R * s_out
            flops = lookups * L * D * D * 2 * C
            dname = str(dt).replace('torch.', '')
            line = {'metric': f'altcorr lookups/s ({name})', 'value': round(lookups / (ms / 1e3), 1),
                    'unit': 'lookups/s', 'workload': name, 'dtype': dname, 'edges': E, 'hw': [h, w], 'radius': r,
                    'path': 'mfma' if getattr(blk, 'mfma', False) else ('fp8 valu' if blk.fp8 else 'per-level'),
                    'ms_per_pass': round(ms, 4), 'pyramid_build_ms': round(build_ms, 4),
                    'GBps_algorithmic': round(nbytes / (ms / 1e3) / 1e9, 1),
                    'hbm_frac': round(nbytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBPS, 4),
                    'TFLOPs_algorithmic': round(flops / (ms / 1e3) / 1e12, 2)}
            if dname in MFMA_PEAK_TFLOPS:
                line['mfma_frac'] = round(flops / (ms / 1e3) / 1e12 / MFMA_PEAK_TFLOPS[dname], 4)
            print(json.dumps(line), flush=True)
            del blk
            torch.cuda.empty_cache()


def window_bench(db, dev, args):
    """The frontend window (12 keyframes at 48x64, 60 neighbour + 10 long edges, 4 GN iterations,
    t0 = 2): the reference's most frequent ba call (hislam2/factor_graph.py:189-233), latency-bound.
    Per call through droid_backends.ba exactly as the reference calls it: host copy of ii/jj,
    planning (cached while the edge list repeats), the iterations on the device.  Reported: the
    median wall-clock latency per call with a synchronisation after it, the device time per call
    (HIP events around back-to-back calls), and both precision modes."""
    import numpy as np
    import torch
    from hislam2_amd import synth
    from hislam2_amd.metrics import reprojection_rms
    p = synth.ba_window()
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    intr, tg, wt, eta, ii, jj = up(p.intrinsics), up(p.targets), up(p.weights), up(p.eta), up(p.ii), up(p.jj)
    line = {'metric': 'frontend-window ba latency', 'workload': 'window: 12 keyframes 48x64, 70 edges, '
            'ba(4 iterations, t0=2, t1=12)', 'edges': len(p.ii), 'iterations': p.iterations}
    rms0 = reprojection_rms(up(p.poses), up(p.disps), intr, tg, wt, ii, jj)
    for name, prec in (('split', 0), ('plain', 1)):
        poses, disps = up(p.poses), up(p.disps)

        def call():
            if prec == 0:
                return db.ba(poses, disps, intr, tg, wt, eta, ii, jj, p.t0, p.t1, p.iterations, p.lm, p.ep, False)
            return db.ba_ex(poses, disps, intr, tg, wt, eta, ii, jj, p.t0, p.t1, p.iterations, p.lm, p.ep, False,
                            prec)

        P_, D_ = up(p.poses), up(p.disps)
        db.ba_ex(P_, D_, intr, tg, wt, eta, ii, jj, p.t0, p.t1, p.iterations, p.lm, p.ep, False, prec)
        rms = reprojection_rms(P_, D_, intr, tg, wt, ii, jj)
        for _ in range(max(args.warmup, 3)):
            call()
        torch.cuda.synchronize()
        lat = []
        for _ in range(20 * args.steps):
            t0 = time.perf_counter()
            call()
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t0)
        dev_ms = time_launches(call, 20 * args.steps, 1)
        line[name] = {'latency_ms_median': round(1e3 * float(np.median(lat)), 4),
                      'device_ms_per_call': round(dev_ms, 4),
                      'GN_it_s_device': round(p.iterations / (dev_ms / 1e3), 1),
                      'reproj_rms_px': round(rms, 6)}
    line['reproj_rms_start_px'] = round(rms0, 6)
    line['speedup_vs_plain_device'] = round(line['plain']['device_ms_per_call'] / line['split']['device_ms_per_call'], 3)
    print(json.dumps(line), flush=True)


def hires64_bench(db, dev, args):
    """The high-resolution workload: 64 keyframes at 96x128 (HW = 12288), 800 edges, ba of 4 GN
    iterations (t0 = 1) plus, per iteration, the radius-4 4-level on-the-fly lookup of all 800
    edges (AltCorrBlock; a stored volume would need 400 MB per edge).  A step = that ba call + 4
    lookups.  Split: bf16 lookup on the matrix cores + FP64 sums and solve; plain: the same
    kernels in FP32.  Also the fp8 (e4m3, per-channel scales) lookup."""
    import numpy as np
    import torch
    from hislam2_amd import corr as hcorr
    from hislam2_amd import synth
    from hislam2_amd.metrics import reprojection_rms
    p = synth.ba_hires64()
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    intr, tg, wt, eta, ii, jj = up(p.intrinsics), up(p.targets), up(p.weights), up(p.eta), up(p.ii), up(p.jj)
    n, h, w = p.disps.shape
    fm = up(synth.fmaps(64, n, h=h, w=w))
    coords = up(synth.lookup_coords(65, len(p.ii), h=h, w=w))
    line = {'metric': 'hires64 GN it/s (ba + on-the-fly lookup)', 'workload': 'hires64: 64 keyframes 96x128, '
            '800 edges, ba(4 iterations, t0=1) + radius-4 4-level AltCorrBlock lookup of all edges per iteration',
            'edges': len(p.ii), 'iterations': p.iterations}
    rms0 = reprojection_rms(up(p.poses), up(p.disps), intr, tg, wt, ii, jj)
    for name, prec, dt in (('split', 0, torch.bfloat16), ('plain', 1, torch.float32)):
        P_, D_ = up(p.poses), up(p.disps)
        out = db.ba_ex(P_, D_, intr, tg, wt, eta, ii, jj, p.t0, p.t1, p.iterations, p.lm, p.ep, False, prec)
        rms = reprojection_rms(P_, D_, intr, tg, wt, ii, jj)
        blk = hcorr.AltCorrBlock(fm, radius=4, dtype=dt)
        poses, disps = up(p.poses), up(p.disps)

        def step():
            db.ba_ex(poses, disps, intr, tg, wt, eta, ii, jj, p.t0, p.t1, p.iterations, p.lm, p.ep, False, prec)
            for _ in range(p.iterations):
                blk(coords, ii, jj)

        ms = time_launches(step, args.steps, args.warmup)
        look = time_launches(lambda: blk(coords, ii, jj), args.steps, 1)
        line[name] = {'GN_it_s': round(p.iterations / (ms / 1e3), 3), 'ms_per_step': round(ms, 3),
                      'lookup_ms': round(look, 4), 'solve_ok': int(out[3].item()), 'reproj_rms_px': round(rms, 6)}
        del blk
        torch.cuda.empty_cache()
    f8 = hcorr.AltCorrBlock(fm, radius=4, dtype=torch.float8_e4m3fn)
    line['fp8_lookup_ms'] = round(time_launches(lambda: f8(coords, ii, jj), args.steps, 1), 4)
    line['reproj_rms_start_px'] = round(rms0, 6)
    line['speedup_vs_plain'] = round(line['split']['GN_it_s'] / line['plain']['GN_it_s'], 3)
    line['reproj_diff_px'] = round(abs(line['split']['reproj_rms_px'] - line['plain']['reproj_rms_px']), 7)
    print(json.dumps(line), flush=True)


def mono_bench(db, p, up, dev, args):
    """The mono-prior step on the global300 graph (300 keyframes, 3000 edges, 48x64, a 2x2
    scale grid per frame): JDSA end to end, and its proj_trans kernel alone (algorithmic bytes:
    targets + weights 16 B per pixel and edge, disparity 4 B and C, w 8 B per pixel and frame)."""
    import numpy as np
    import torch
    from hislam2_amd.geom import ba as gba
    rng = np.random.default_rng(3)
    n, h, w = p.disps.shape
    prior = p.disps_true / (1.0 + 0.1 * rng.uniform(-1, 1, size=(n, 1, 1)))
    prior[rng.uniform(size=prior.shape) < 0.1] = 0.0
    ii, jj = up(p.ii), up(p.jj)
    M = len(np.unique(p.ii))
    poses, disps = up(p.poses), up(p.disps)[None]
    intr = up(p.intrinsics)[None, None].expand(1, n, 4)
    tg, wt = up(p.targets), up(p.weights)
    eta = torch.full((M, h, w), 1e-4, device=dev)
    prior_t = up(prior.astype(np.float32))
    scales0 = up((1.0 + 0.05 * rng.normal(size=(n, 2, 2))).astype(np.float32))

    def jdsa():
        return gba.JDSA(tg, wt, eta, poses[None], disps, intr, prior_t, scales0.clone(), ii, jj, 0.01)

    def pt():
        return db.proj_trans(poses, disps[0], intr[0, 0].contiguous(), tg, wt, ii, jj)

    res = {}
    for name, fn in (('jdsa', jdsa), ('proj_trans', pt)):
        for _ in range(max(args.warmup, 1)):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        res[name] = e0.elapsed_time(e1) / args.steps
    HW = h * w
    pt_bytes = len(p.ii) * HW * 16 + M * HW * 12
    print(json.dumps({'metric': 'mono-prior step (JDSA) on global300', 'jdsa_ms': round(res['jdsa'], 4),
                      'proj_trans_ms': round(res['proj_trans'], 4),
                      'proj_trans_GBps': round(pt_bytes / (res['proj_trans'] / 1e3) / 1e9, 1),
                      'proj_trans_bytes': pt_bytes, 'frames': M, 'edges': len(p.ii),
                      'note': 'proj_trans_ms includes the host-side edge grouping and its upload'}), flush=True)


def pgba_bench(db, dev, args):
    """One Sim3 pose-graph BA step (droid_backends.pgba) on a loop-closure graph of 300 keyframes
    at 48x64: poses 8..299 free (292 x 7 parameters), 638 image edges (neighbours plus 40 long
    edges), 320 pose-graph pairs.  The timed call includes the host-side planning, the upload of
    its tables and the final synchronisation, as every call of the reference plans on the host."""
    import numpy as np
    import torch
    from hislam2_amd import synth
    q = synth.pgba_problem(seed=5, n=300, radius=1, n_long=40, t0=8, n_pg=320)
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in q.items() if isinstance(v, np.ndarray)}
    disps0 = t['disps'].clone()

    def call():
        return db.pgba(t['poses'], t['disps'], t['intrinsics'], t['targets'], t['weights'], t['eta'], t['Hsp'],
                       t['vsp'], t['ii'], t['jj'], t['iip'], t['jjp'], q['t0'], q['t1'], q['lm'], q['ep'])

    for _ in range(max(args.warmup, 1)):
        call()
        t['disps'].copy_(disps0)
    torch.cuda.synchronize()
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        call()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        t['disps'].copy_(disps0)
    ms = 1e3 * float(np.median(times))
    line = {'metric': 'pgba step on loop300', 'pgba_ms': round(ms, 4), 'poses': q['t1'] - q['t0'],
            'edges': len(q['ii']), 'pose_graph_pairs': len(q['iip']), 'hw': list(q['disps'].shape[1:])}
    if not args.no_cpu_baseline:
        from oracle import pgba as opg
        c0 = time.perf_counter()
        opg.pgba(q['poses'], q['disps'], q['intrinsics'], q['targets'], q['weights'], q['eta'], q['Hsp'], q['vsp'],
                 q['ii'], q['jj'], q['iip'], q['jjp'], q['t0'], q['t1'], q['lm'], q['ep'])
        line['cpu_baseline'] = {'value': round(1e3 * (time.perf_counter() - c0), 1), 'unit': 'ms per step',
                                'cores': 1, 'kind': 'port', 'sample': 'one full step of the same graph'}
    print(json.dumps(line), flush=True)


def geometry_bench(db, p, up, args):
    """The edge-selection kernels on the global300 graph: frame_distance of all 3000 edges,
    depth_filter of all 300 frames (six neighbours each), iproj of all 300 frames."""
    import numpy as np
    import torch
    poses, disps, intr = up(p.poses), up(p.disps), up(p.intrinsics)
    ii, jj = up(p.ii), up(p.jj)
    n = p.poses.shape[0]
    inds = torch.arange(n, device=disps.device)
    thresh = torch.full((n,), 0.05, device=disps.device)
    calls = {'frame_distance': lambda: db.frame_distance(poses, disps, intr, ii, jj, 0.7),
             'depth_filter': lambda: db.depth_filter(poses, disps, intr, inds, thresh),
             'iproj': lambda: db.iproj(poses, disps, intr)}
    res = {}
    for name, fn in calls.items():
        for _ in range(max(args.warmup, 1)):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        res[name + '_ms'] = round(e0.elapsed_time(e1) / args.steps, 4)
    print(json.dumps({'metric': 'edge-selection geometry on global300', **res, 'edges': len(p.ii), 'frames': n,
                      'note': 'frame_distance_ms and depth_filter_ms include the host-side index checks'}),
          flush=True)


GOLD = os.path.join(ROOT, 'tests', 'golden')


def ba_oracle_errors(poses, disps, dx, dz):
    """Errors of the 8-iteration global300 ba call against the float64 oracle's result stored in
    tests/golden/ba_global300_8_iterations.npz: final poses and disparities relative to their
    largest entry; the last dx and dz entry by entry relative to their own magnitude plus the
    first iteration's largest update."""
    import numpy as np
    path = os.path.join(GOLD, 'ba_global300_8_iterations.npz')
    if not os.path.exists(path):
        return None
    f = np.load(path)
    rel = lambda g, r: float(np.max(np.abs(g - r)) / np.max(np.abs(r)))
    loc = lambda g, r, fl: float(np.max(np.abs(g - r) / (np.abs(r) + fl)))
    return {'poses': rel(poses.double().cpu().numpy(), f['out_poses']),
            'disps': rel(disps.double().cpu().numpy().reshape(-1)[f['disps_index']], f['out_disps']),
            'dx_last': loc(dx.double().cpu().numpy(), f['out_dx'], np.max(np.abs(f['out_dx_it1']))),
            'dz_last': loc(dz.double().cpu().numpy().reshape(-1)[f['dz_index']], f['out_dz'], float(f['dz1_absmax']))}


def lookup_oracle_errors(out):
    """Max-norm relative and absolute error of a [1, 3000, 196, 48, 64] global300 lookup at the
    edges and pixels of tests/golden/corrblock_global300_bf16.npz (float64 oracle)."""
    import numpy as np
    import torch
    path = os.path.join(GOLD, 'corrblock_global300_bf16.npz')
    if not os.path.exists(path):
        return None
    f = np.load(path)
    e = torch.from_numpy(f['edges']).to(out.device)
    o = out[0].index_select(0, e).flatten(2)
    px = torch.from_numpy(f['pixels']).to(out.device)
    got = torch.gather(o, 2, px[:, None, :].expand(-1, o.shape[1], -1)).double().cpu().numpy()
    d = np.abs(got - f['out'])
    return {'rel': float(d.max() / np.abs(f['out']).max()), 'abs': float(d.max())}


def lookup_traffic():
    """HBM bytes of one bf16 lookup launch, from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against a 2 GiB device copy); None if absent."""
    path = os.path.join(ROOT, 'profiles', 'corr_lookup_bf16_global300_pmc.md')
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for ln in fh:
            if ln.startswith('bytes_per_launch'):
                return int(ln.split(':')[1])
    return None


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as fh:
            for ln in fh:
                if ln.startswith('model name'):
                    return ln.split(':', 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(p, fm32, coords):
    """The reference's algorithm on the host (oracle/torch_cpu.py: ba_cuda's semantics in torch
    float32 with index_add_ segment sums and an FP64 Cholesky; CorrBlock as matmul + avg_pool2d +
    grid_sample), on all host threads: one warm-up, then the median of 5 GN iterations at each of
    the 4-keyframe (C1), frontend-window (C2) and global300 (C3) problems, and the lookup of 10 of
    the 3000 global300 edges (median of 3) scaled to all of them.  value = GN iterations/s of the
    global300 step (one GN iteration + the lookup of all edges).  The float64 numpy oracle's
    global300 GN iteration is timed once beside it."""
    import numpy as np
    import torch
    from hislam2_amd import synth
    sys.path.insert(0, ROOT)
    from oracle import ba as oba
    from oracle import torch_cpu as tcpu
    torch.set_num_threads(os.cpu_count() or 1)
    threads = torch.get_num_threads()

    def median_time(fn, reps):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    ba_s = {}
    for name in ('small', 'window', 'global300'):
        q = p if name == 'global300' else synth.PROBLEMS[name]()
        log('cpu baseline: torch GN iteration,', name)
        ba_s[name] = median_time(lambda: tcpu.ba(q.poses, q.disps, q.intrinsics, q.targets, q.weights, q.eta, q.ii,
                                                 q.jj, q.t0, q.t1, 1, q.lm, q.ep), 5)
    n = 10
    fm = fm32[0].float().cpu()
    f1, f2 = fm[p.ii[:n]][None], fm[p.jj[:n]][None]
    c = coords[:, :n].float().cpu()
    log('cpu baseline: torch lookup of', n, 'edges')
    lk = median_time(lambda: tcpu.corrblock_lookup(f1, f2, c), 3)
    per_iter = ba_s['global300'] + lk * len(p.ii) / n
    log('cpu baseline: float64 oracle GN iteration (global300)')
    t0 = time.perf_counter()
    oba.ba(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.eta, p.ii, p.jj, p.t0, p.t1, 1, p.lm, p.ep)
    t_or = time.perf_counter() - t0
    return {'value': round(1.0 / per_iter, 6), 'unit': 'GN it/s', 'cores': threads, 'kind': 'port',
            'sample': f'torch float32 on {threads} host threads ({_cpu_model()}): median of 5 global300 GN '
                      f'iterations ({ba_s["global300"]:.3f} s) + CorrBlock lookup of {n}/3000 edges '
                      f'(median of 3, {lk:.3f} s, scaled x{len(p.ii) // n})',
            'ba_ms_per_iteration': {k: round(1e3 * v, 3) for k, v in ba_s.items()},
            'lookup_ms_per_edge': round(1e3 * lk / n, 3),
            'oracle_float64_global300_iteration_s': round(t_or, 2)}


if __name__ == '__main__':
    main()
