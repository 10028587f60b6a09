#!/usr/bin/env python3
"""Benchmark of the hot path: factorized backprojection of a 12,288 x 8,192 spotlight image from 16,384 pulses x
8,192 samples (complex64) on one MI355X, with the error against a float64 exact backprojection on 16 seeded
256 x 256 tiles, the dominant kernel's roofline fraction, exact backprojection of the same history (the whole grid and
a 4,096 x 4,096 grid), and the reference's CPU former timed on the whole image on the host's cores.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--T 8]

A step is one image formation (ImageFormer call) with the phase history already resident in device memory and
the image written to device memory; the window and the scan for NaN and inf are part of it. The data are
synthetic: 2,000 point targets with Rayleigh amplitudes uniform over the 6 km x 4 km scene plus white clutter
25 dB below the targets' power, X band (9.6 GHz, 300 MHz), 10 km slant range, 30 degrees grazing, 6 degree
aperture. Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 's/image + Gpixel·pulses/s at 12288×8192, at ≤−60 dB error vs FP64 oracle'
# (final tile, levels) the batch's autotuner chooses among: four levels and tiles of 8 or 4 (three-level plans and
# tiles of 16 missed -60 dB on some patches by up to 2.5 dB at predicted errors of -63 to -69 dB)
BATCH_CANDIDATES = ((8, 4), (4, 4))
HBM_PEAK = 8.0e12             # bytes/s
FP32_MATRIX_PEAK = 157.3e12   # flop/s, MI355X dense float32 matrix rate


def log(*a):
    print('[bench]', *a, file=sys.stderr, flush=True)


def workload(small=False):
    """(col, grid, targets, clutter_db, seed): the 12,288 x 8,192 image from 16,384 x 8,192 samples, or (small) the
    512 x 512 image from 512 x 1,024 samples."""
    from fastsar_amd import sim
    if small:
        col = sim.spotlight(512, 1024, 3.0)
        grid = dict(nx=512, ny=512, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
        return col, grid, 25, 256.0, 256.0, -30.0, 1
    col = sim.spotlight(16384, 8192, 6.0)
    sp = 6000.0 / 12288
    grid = dict(nx=12288, ny=8192, spx=sp, spy=sp, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    return col, grid, 2000, 6000.0, 4000.0, -25.0, 2


def accuracy_points(grid, ntiles, side, seed):
    rng = np.random.default_rng(seed)
    nx, ny = grid['nx'], grid['ny']
    side = min(side, nx, ny)
    e1, e2 = np.asarray(grid['e1']), np.asarray(grid['e2'])
    tiles, pts = [], []
    for _ in range(ntiles):
        i0, j0 = int(rng.integers(0, nx - side + 1)), int(rng.integers(0, ny - side + 1))
        tiles.append((i0, j0))
        I, J = np.meshgrid(np.arange(i0, i0 + side), np.arange(j0, j0 + side), indexing='ij')
        p = ((I - nx / 2.0) * grid['spx'])[..., None] * e1 + ((J - ny / 2.0) * grid['spy'])[..., None] * e2
        pts.append(p.reshape(-1, 3))
    return tiles, np.concatenate(pts), side


def target_tiles(grid, targets, side):
    """Tiles of side^2 pixels, each centred (as far as the grid allows) on the pixel nearest a target."""
    nx, ny = grid['nx'], grid['ny']
    side = min(side, nx, ny)
    e1, e2 = np.asarray(grid['e1']), np.asarray(grid['e2'])
    tiles, pts = [], []
    for t in np.asarray(targets, np.float64):
        ci = int(round(t @ e1 / grid['spx'] + nx / 2.0))
        cj = int(round(t @ e2 / grid['spy'] + ny / 2.0))
        i0 = min(max(ci - side // 2, 0), nx - side)
        j0 = min(max(cj - side // 2, 0), ny - side)
        tiles.append((i0, j0))
        I, J = np.meshgrid(np.arange(i0, i0 + side), np.arange(j0, j0 + side), indexing='ij')
        p = ((I - nx / 2.0) * grid['spx'])[..., None] * e1 + ((J - ny / 2.0) * grid['spy'])[..., None] * e2
        pts.append(p.reshape(-1, 3))
    return tiles, np.concatenate(pts), side


def cpu_model():
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline(col, grid, S_host, images, tiles, side, ref_vals, T_cpu=16):
    """The reference's CPU former (its own compiled kernels, fastsar/ffbp_cpu.cpp, driven as fastsar/ffbp_cpu.py:158-291
    drives it) on the whole image, host to host (window pass, then every first-level child in groups of 8): after an
    untimed warm call on one child, the median of `images` full images, and the error of the last one on the accuracy
    tiles against the float64 checker."""
    import oracle
    from fastsar_amd import plan as pl
    from fastsar_amd.api import _window
    from fastsar_amd.metrics import rel_db
    os.environ.setdefault('OMP_PLACES', 'cores')
    os.environ.setdefault('OMP_PROC_BIND', 'close')
    L = oracle.ref_lib()
    threads = int(L.ffbp_cpu_threads())
    p = pl.make_plan(col, grid['nx'], grid['ny'], grid['spx'], grid['spy'], T=T_cpu, nlev=3, e1=grid['e1'], e2=grid['e2'])
    coll = pl.collection_arrays(p, col.ant)
    former = oracle.RefCpuFormer(p, coll)
    G = p['levels'][0]['C']
    wp, wk = _window(col.Np, col.K)
    former(S_host[:, :], children=[G - 1], ng=1)                   # warm: one child, untimed
    times, img = [], None
    for _ in range(max(1, images)):
        t0 = time.perf_counter()
        img = former(former.window(S_host, wp, wk), ng=8)
        times.append(time.perf_counter() - t0)
        log(f'cpu baseline: one image in {times[-1]:.2f} s')
    t = float(np.median(times))
    got = np.concatenate([img[i0:i0 + side, j0:j0 + side].ravel() for i0, j0 in tiles])
    err = rel_db(got, ref_vals)
    pairs = grid['nx'] * grid['ny'] * col.Np
    return dict(value=round(pairs / t / 1e9, 3), unit='Gpixel·pulses/s', cores=threads, kind='reference',
                sample=(f'reference CPU former (fastsar/ffbp_cpu.cpp built from its sources, x86-64-{L.isa}, OpenMP '
                        f'{threads} threads on {cpu_model()}), T={T_cpu}, 3 levels: the whole image host to host, '
                        f'median of {len(times)} after a warm call on one first-level child, no extrapolation'),
                s_per_image=round(t, 3), times_s=[round(v, 3) for v in times], error_db=round(err, 2),
                error_tiles=f'{len(tiles)} x {side}^2', cpu_model=cpu_model(), threads=threads)


def exact_leg(col, grid, dS, tiles, side, ref_vals, steps, full=True):
    """Exact backprojection (ExactFormer) of the same history: the whole grid with cubic reads (time of one warm call
    and its error on the accuracy tiles against the float64 checker; skipped with full=False), and a 4,096 x 4,096 grid
    about the scene centre with cubic and with linear reads (median of `steps` calls). Pixel-pulse pairs per second."""
    out = {}
    if full:
        out.update(_exact_full(col, grid, dS, tiles, side, ref_vals))
    _exact_4096(col, grid, dS, steps, out)
    return out


def _exact_full(col, grid, dS, tiles, side, ref_vals):
    from fastsar_amd import _hip
    from fastsar_amd.exact import ExactFormer
    from fastsar_amd.metrics import rel_db
    out = {}
    f = ExactFormer(col.ant, col.fmin, col.df, col.K, **grid)
    img = _hip.DeviceArray((grid['nx'], grid['ny']), np.complex64)
    f(dS, out=img)
    _hip.check(_hip.lib().fsh_synchronize())
    t0 = time.perf_counter()
    f(dS, out=img)
    _hip.check(_hip.lib().fsh_synchronize())
    t = time.perf_counter() - t0
    h = img.to_host()
    err = rel_db(np.concatenate([h[i0:i0 + side, j0:j0 + side].ravel() for i0, j0 in tiles]), ref_vals)
    pairs = grid['nx'] * grid['ny'] * col.Np
    out['full'] = dict(interp='cubic', tile=list(f.tile), ms=round(t * 1e3, 1), gpairs_per_s=round(pairs / t / 1e9, 1),
                       error_db=round(err, 2), error_tiles=f'{len(tiles)} x {side}^2')
    log(f'exact backprojection {grid["nx"]}x{grid["ny"]} from {col.Np} x {col.K}, cubic: {out["full"]}')
    return out


def _exact_4096(col, grid, dS, steps, out):
    from fastsar_amd import _hip
    from fastsar_amd.exact import ExactFormer
    g4 = dict(grid, nx=4096, ny=4096)
    img = _hip.DeviceArray((4096, 4096), np.complex64)
    for interp in ('cubic', 'linear'):
        f = ExactFormer(col.ant, col.fmin, col.df, col.K, **g4, interp=interp)
        f(dS, out=img)
        ts = []
        for _ in range(max(1, steps)):
            _hip.check(_hip.lib().fsh_synchronize())
            t0 = time.perf_counter()
            f(dS, out=img)
            _hip.check(_hip.lib().fsh_synchronize())
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        pairs = 4096 * 4096 * col.Np
        out[f'4096x4096_{interp}'] = dict(ms=round(t * 1e3, 2), gpairs_per_s=round(pairs / t / 1e9, 1))
        log(f'exact backprojection 4096x4096 from {col.Np} x {col.K}, {interp}: {out[f"4096x4096_{interp}"]}')
        del f


def measured_traffic(label, former, args):
    """HBM bytes per launch of `label` from the committed counter summary (FETCH_SIZE + WRITE_SIZE scaled by a copy
    kernel's reading), when the summary covers this workload; else None."""
    path = os.path.join(ROOT, 'profiles', f'ffbp_12288x8192_T{former.T}_hbm_traffic.md')
    if (args.small or former.levels != 4 or args.atten != 80.0 or args.taps or args.train_taps or
            args.history != 'complex64' or args.precision != 'float32' or not os.path.exists(path)):
        return None
    for line in open(path):
        if line.startswith('traffic_bytes_per_launch '):
            return json.loads(line.split(' ', 1)[1]).get(label)
    return None


def bench_batch(args):
    """The batched workload: 48 images of uneven size (256^2 to 4096^2 px, 512 to 8192 pulses x 1024 to 4096
    samples, spotlight and 2-degree-squint patches), each former's final tile and levels chosen by the autotuner,
    formed by one hipGraph launch; against the same images formed by one former call each."""
    out = run_batch(args, tune=True)
    print(json.dumps(out), flush=True)


def run_batch(args, tune):
    """The batch of bench_batch -> its result line (dict). tune=False: each image's final tile chosen by its predicted
    error alone (T 8, else 4, four levels), no timing of candidates (the default run's bounded form)."""
    from fastsar_amd import ImageFormer, _hip, plan as pl, sim
    from fastsar_amd.batch import Batch, autotune
    from fastsar_amd.exact import simulate_device, bp_points_f64_device
    from fastsar_amd.metrics import rel_db
    from fastsar_amd.api import _window
    scenes = sim.uneven_batch(seed=5, n=args.batch_images)
    formers, hists = [], []
    t0 = time.perf_counter()
    for i, sc in enumerate(scenes):
        col, g = sc['col'], sc['grid']
        dS = simulate_device(col, sc['pos'], sc['amp'], clutter_sigma=sc['sigma'], seed=100 + i)
        if tune:
            f = autotune(col.ant, col.fmin, col.df, col.K, **g, S=dS, target_db=args.batch_target, atten=args.batch_atten,
                         candidates=BATCH_CANDIDATES)
        else:
            T, _ = pl.choose_T(col.ant, col.fmin + col.K * col.df, g['nx'], g['ny'], g['spx'], g['spy'], g['e1'], g['e2'],
                               args.batch_target, choices=(8, 4))
            f = ImageFormer(col.ant, col.fmin, col.df, col.K, **g, T=T, levels=4, atten=args.batch_atten)
            f.tuned = dict(T=T, levels=4, ms=None, tried=0)
        formers.append(f)
        hists.append(dS)
        log(f'image {i}: {g["nx"]}^2 from {col.Np} x {col.K}, squint {sc["squint"]:g}: T={f.T}, {f.levels} levels '
            f'({f.tuned["tried"]} tried, {f.tuned["ms"] and round(f.tuned["ms"], 2)} ms)')
    log(f'histories and formers in {time.perf_counter() - t0:.1f} s')
    b = Batch(formers, hists, streams=args.batch_streams)
    for _ in range(args.warmup):
        b.run()
        b.run_loop()
    runs = [b.run() for _ in range(args.steps)]
    loops = [b.run_loop() for _ in range(args.steps)]
    ms = float(np.median([r['ms'] for r in runs]))
    lms = float(np.median([r['ms'] for r in loops]))
    done = np.median(np.stack([r['done_ms'] for r in runs]), axis=0)
    each = np.median(np.stack([r['each_ms'] for r in loops]), axis=0)
    pairs = sum(sc['grid']['nx'] * sc['grid']['ny'] * sc['col'].Np for sc in scenes)
    log(f'batch: {ms:.1f} ms ({pairs / ms / 1e6:.1f} Gpixel.pulses/s); loop of single calls: {lms:.1f} ms '
        f'({pairs / lms / 1e6:.1f})')
    # error of every image against the float64 checker on two 64^2 tiles, each about one of the image's targets (a
    # tile of clutter alone measures the error of the bright targets' leakage against a background 30 dB down)
    errs = []
    for i, (sc, f, dS) in enumerate(zip(scenes, formers, hists)):
        col, g = sc['col'], sc['grid']
        tiles, pts, side = target_tiles(g, sc['pos'][:2], 64)
        nfft = 1 << int(np.ceil(np.log2(args.upsample * col.K)))
        wp, wk = _window(col.Np, col.K)
        ref = bp_points_f64_device(dS, col.ant, np.linalg.norm(col.ant, axis=1), col.fmin, col.df, pts, nfft, True,
                                   wp, wk, chunk=max(1, (1 << 30) // (nfft * 16)))
        img = b.images[i].to_host()
        got = np.concatenate([img[i0:i0 + side, j0:j0 + side].ravel() for i0, j0 in tiles])
        errs.append(rel_db(got, ref))
        log(f'image {i}: {g["nx"]}^2, {col.Np} x {col.K}, band {col.K * col.df / 1e6:.0f} MHz, aperture '
            f'{np.ptp(col.ant[:, 1]):.0f} m, squint {sc["squint"]:g}, T={f.T}/{f.levels}, predicted '
            f'{f.predicted_error_db:.1f} dB: {errs[-1]:.1f} dB')
    errs = np.array(errs)
    log(f'error vs float64 exact backprojection: worst {errs.max():.1f} dB, median {np.median(errs):.1f} dB')
    choice = {}
    for f in formers:
        choice[f'T{f.T}/L{f.levels}'] = choice.get(f'T{f.T}/L{f.levels}', 0) + 1
    del b, formers, hists
    out = dict(metric='aggregate Gpixel·pulses/s, 48 uneven images in one hipGraph launch', value=round(pairs / ms / 1e6, 2),
               unit='Gpixel·pulses/s', n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=round(ms, 2),
               higher_is_better=True, scaling='weak', vs_baseline=None, dtype='f32', data='synthetic',
               config=dict(workload=(f'{len(scenes)} images, sides 256-4096 px log-uniform (seed 5), 512-8192 pulses x '
                                     '1024-4096 samples, spotlight and 2-degree-squint patches, '
                                     + ('autotuned T and levels, ' if tune else 'T by predicted error, 4 levels, ')
                                     + f'{args.batch_streams} streams in one graph'),
                           pixel_pulses=pairs, streams=args.batch_streams,
                           latency_ms_batched=dict(p50=round(float(np.percentile(done, 50)), 2),
                                                   p99=round(float(np.percentile(done, 99)), 2)),
                           latency_ms_single=dict(p50=round(float(np.percentile(each, 50)), 2),
                                                  p99=round(float(np.percentile(each, 99)), 2)),
                           loop_ms=round(lms, 2), loop_value=round(pairs / lms / 1e6, 2), speedup_vs_loop=round(lms / ms, 3),
                           error_db_worst=round(float(errs.max()), 2), error_db_median=round(float(np.median(errs)), 2),
                           choice=choice, target_db=args.batch_target, atten_db=args.batch_atten))
    return out


def bench_learned(args):
    """The learned decimation filters on the small spotlight geometry at 1024 pulses (512^2 px): image error against
    the FP64 checker on held-out scenes and time per image, for the plan's Kaiser filters (70 and 80 dB), the learned
    filters of 6 and 8 taps per phase (the committed ones, trained on 200 scenes) and the same-length Kaiser
    fallback they correct."""
    from fastsar_amd import ImageFormer, _hip, learned
    from fastsar_amd.exact import bp_points_f64_device
    from fastsar_amd.metrics import rel_db
    from fastsar_amd.api import _window
    col, grid, plan = learned.c1_geometry()
    wp, wk = _window(col.Np, col.K)
    pts = np.stack(np.meshgrid(np.arange(grid['nx']), np.arange(grid['ny']), indexing='ij'), -1).reshape(-1, 2)
    e1, e2 = np.asarray(grid['e1']), np.asarray(grid['e2'])
    xyz = ((pts[:, 0] - grid['nx'] / 2.0) * grid['spx'])[:, None] * e1 + ((pts[:, 1] - grid['ny'] / 2.0) * grid['spy'])[:, None] * e2
    args_f = (col.ant, col.fmin, col.df, col.K)
    kw = dict(T=16, levels=3)
    forms = dict(kaiser70=ImageFormer(*args_f, **grid, **kw), kaiser80=ImageFormer(*args_f, **grid, **kw, atten=80.0))
    taps_of = {}
    for T in (6, 8):
        t = learned.default_taps(T)
        if t is None:
            log(f'no committed filters of {T} taps per phase; training on 200 scenes')
            t = learned.train_c1(T, 200, log=log)
        taps_of[T] = t
        forms[f'learned{T}'] = ImageFormer(*args_f, **grid, **kw, taps=t)
        forms[f'fallback{T}'] = ImageFormer(*args_f, **grid, **kw, taps=learned.LearnedTaps(T, 3, {}))
    res = {k: dict(err=[], ms=[]) for k in forms}
    img = _hip.DeviceArray((grid['nx'], grid['ny']), np.complex64)
    for seed in range(5000, 5000 + args.learned_scenes):
        dS = learned.c1_scene(col, seed)
        nfft = 1 << int(np.ceil(np.log2(64 * col.K)))
        ref = bp_points_f64_device(dS, col.ant, np.linalg.norm(col.ant, axis=1), col.fmin, col.df, xyz, nfft, True, wp, wk,
                                   chunk=64)
        for k, f in forms.items():
            f(dS, out=img)
            _hip.check(_hip.lib().fsh_synchronize())
            t0 = time.perf_counter()
            for _ in range(args.steps):
                f(dS, out=img)
            _hip.check(_hip.lib().fsh_synchronize())
            res[k]['ms'].append((time.perf_counter() - t0) / args.steps * 1e3)
            res[k]['err'].append(rel_db(img.to_host().ravel(), ref))
    rows = {}
    for k, f in forms.items():
        lv = f._plan['levels']
        rows[k] = dict(error_db=round(float(np.mean(res[k]['err'])), 2), ms=round(float(np.median(res[k]['ms'])), 3),
                       range_taps=[l['fir_k']['L'] if l['fir_k'] else 1 for l in lv],
                       pulse_taps=[l['fir_p']['L'] if l['fir_p'] else 1 for l in lv])
        log(f'{k:10s} {rows[k]}')
    out = dict(metric='learned decimators: image error (dB vs FP64) and ms per 512^2 image from 1024 x 1024', value=rows['learned6']['error_db'],
               unit='dB', n_gpus=1, steps=args.steps, warmup=1, ms_per_step=rows['learned6']['ms'], higher_is_better=False,
               scaling='weak', vs_baseline=None, dtype='f32', data='synthetic',
               config=dict(workload=f'512x512 from 1024 pulses x 1024 samples, T=16, 3 levels, {args.learned_scenes} held-out scenes',
                           forms=rows, training={T: t.history for T, t in taps_of.items()}))
    print(json.dumps(out, default=float), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--T', type=int, default=8)
    ap.add_argument('--atten', type=float, default=80.0, help='stopband attenuation of the decimation filters (dB)')
    ap.add_argument('--levels', type=int, default=4, help='levels of the plan (the former API default is 3, more when 3 cannot split)')
    ap.add_argument('--tiles', type=int, default=16, help='accuracy tiles (256 x 256) against float64 backprojection')
    ap.add_argument('--upsample', type=int, default=64, help='oversampling of the float64 checker profiles')
    ap.add_argument('--cpu-images', type=int, default=1, help='whole images the reference CPU former forms (median)')
    ap.add_argument('--host-steps', type=int, default=5, help='also time calls from a host history to a host image (median)')
    ap.add_argument('--no-exact', action='store_true', help='skip the exact-backprojection leg')
    ap.add_argument('--no-batch', action='store_true', help='skip the 48-image batch in the default run')
    ap.add_argument('--exact', action='store_true', help='the exact-backprojection leg alone (for profiles)')
    ap.add_argument('--no-full', action='store_true', help='with --exact: the 4096 x 4096 grids only')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-profile', action='store_true', help='timed steps only (for counter runs under rocprofv3)')
    ap.add_argument('--calibrate-copy', action='store_true', help='also run the plain copy kernel on 1 GiB three times')
    ap.add_argument('--small', action='store_true', help='the 512 x 512 scene instead (quick check)')
    ap.add_argument('--learned', action='store_true', help='the learned decimation filters against the Kaiser ones')
    ap.add_argument('--learned-scenes', type=int, default=4)
    ap.add_argument('--batch', action='store_true', help='the batched workload of uneven images instead')
    ap.add_argument('--batch-images', type=int, default=48)
    ap.add_argument('--batch-streams', type=int, default=4)
    ap.add_argument('--batch-target', type=float, default=-64.0, help='predicted final-stage error the autotuner meets (dB; 4 dB below the -60 dB goal, the prediction reading up to 4 dB low on squinted patches)')
    ap.add_argument('--batch-atten', type=float, default=80.0, help='stopband attenuation of the batch formers\' filters (dB)')
    ap.add_argument('--taps', default=None, help='learned decimation filters (an .npz of fastsar_amd.learned) for the image')
    ap.add_argument('--no-learned', action='store_true', help='skip the learned form beside the plain one')
    ap.add_argument('--no-valu-form', action='store_true', help='skip the float32 VALU form beside the matrix-core one')
    ap.add_argument('--train-taps', type=int, default=0, help='train filters of this many taps per phase on 2 scenes of '
                    'the workload first (seeds 11, 12), saved to --taps-out, and form with them')
    ap.add_argument('--taps-out', default=None, help='where --train-taps saves its filters (default learned_T<n>.npz)')
    ap.add_argument('--precision', default='float32', choices=['float32', 'float16'],
                    help='float16: the two-pass levels store their intermediate rows in float16')
    ap.add_argument('--history', default='complex64', choices=['complex64', 'float16', 'bfloat16'],
                    help='storage of the phase history the former reads (float16: per-pulse power-of-two scaled)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU: --gpus {args.gpus} is not supported (the image-formation path needs no '
              'collective: a sharded former would give each GPU a contiguous range of first-level children and gather '
              'the image blocks)', file=sys.stderr)
        sys.exit(2)

    if args.batch:
        return bench_batch(args)
    if args.learned:
        return bench_learned(args)
    from fastsar_amd import ImageFormer, _hip, sim
    from fastsar_amd.exact import simulate_device, bp_points_f64_device
    from fastsar_amd.metrics import rel_db, stage_bytes, stage_flops
    from fastsar_amd.api import _window

    dev = _hip.device_info()
    log('device', dev)
    col, grid, ntg, ext1, ext2, cl_db, seed = workload(args.small)
    rng = np.random.default_rng(seed)
    pos, amp = sim.point_scene(rng, ntg, ext1, ext2, grid['e1'], grid['e2'])
    t0 = time.perf_counter()
    dS = simulate_device(col, pos, amp, clutter_sigma=sim.clutter_sigma(amp, cl_db), seed=seed)
    log(f'phase history {col.Np} x {col.K} generated on the device in {time.perf_counter() - t0:.2f} s')
    if args.exact:
        tiles, pts, side, ref_vals = [], None, 0, None
        if not args.no_full:
            tiles, pts, side = accuracy_points(grid, args.tiles, 256, seed=7)
            nfft = 1 << int(np.ceil(np.log2(args.upsample * col.K)))
            wp, wk = _window(col.Np, col.K)
            ref_vals = bp_points_f64_device(dS, col.ant, np.linalg.norm(col.ant, axis=1), col.fmin, col.df, pts, nfft, True,
                                            wp, wk, chunk=max(1, (1 << 30) // (nfft * 16)))
        ex = exact_leg(col, grid, dS, tiles, side, ref_vals, args.steps, full=not args.no_full)
        lin = ex['4096x4096_linear']
        print(json.dumps(dict(metric='exact backprojection Gpixel·pulses/s, 4096x4096 from 16384 x 8192, linear reads',
                              value=lin['gpairs_per_s'], unit='Gpixel·pulses/s', n_gpus=1, steps=args.steps, warmup=1,
                              ms_per_step=lin['ms'], higher_is_better=True, scaling='weak', vs_baseline=None, dtype='f32',
                              data='synthetic', config=dict(workload='exact backprojection on the 12288x8192 benchmark '
                                                                     'geometry', exact_form=ex))), flush=True)
        return
    hist, hist_host = dS, None
    if args.history != 'complex64':
        from fastsar_amd import precision as pr
        hist_host = pr.pack(dS.to_host(), args.history)
        hist = hist_host.to_device()
        log(f'history stored as {args.history} pairs ({hist_host.data.nbytes / 1e9:.2f} GB)')

    taps = None
    if args.train_taps:
        from fastsar_amd import learned, plan as pl
        p0 = pl.make_plan(col, grid['nx'], grid['ny'], grid['spx'], grid['spy'], T=args.T, nlev=args.levels, atten=args.atten,
                          e1=grid['e1'], e2=grid['e2'])
        c0 = pl.collection_arrays(p0, col.ant)

        def scenes():
            for sd in (11, 12):
                r = np.random.default_rng(sd)
                ps, am = sim.point_scene(r, ntg, ext1, ext2, grid['e1'], grid['e2'])
                yield simulate_device(col, ps, am, clutter_sigma=sim.clutter_sigma(am, cl_db), seed=sd)
        t0 = time.perf_counter()
        taps = learned.train_scenes(p0, c0, scenes(), args.train_taps, window=_window(col.Np, col.K), iters=6000, log=log)
        log(f'trained {args.train_taps} taps per phase in {time.perf_counter() - t0:.1f} s: {taps.history}')
        out = args.taps_out or f'learned_T{args.train_taps}.npz'
        if os.path.dirname(out):
            os.makedirs(os.path.dirname(out), exist_ok=True)
        taps.save(out)
    elif args.taps:
        from fastsar_amd import learned
        taps = learned.LearnedTaps.load(args.taps)
    t0 = time.perf_counter()
    former = ImageFormer(col.ant, col.fmin, col.df, col.K, **grid, T=args.T, levels=args.levels or None,
                         atten=args.atten, taps=taps, precision=args.precision)
    plan = former._plan
    form_T, form_levels, form_pred = former.T, former.levels, former.predicted_error_db
    log(f'former built in {time.perf_counter() - t0:.2f} s: T={former.T}, levels={former.levels}, '
        f'splits={plan["split"]}, final {plan["final"]["P"]} x {plan["final"]["K"]}, predicted {former.predicted_error_db:.1f} dB')
    for l, lv in enumerate(plan['levels']):
        log(f'  level {l}: {lv["sx"]}x{lv["sy"]} children, Dk={lv["Dk"]} Dp={lv["Dp"]}, [{lv["P"]} x {lv["K"]}] -> '
            f'[{lv["Po"]} x {lv["Ko"]}], Lk={lv["fir_k"]["L"] if lv["fir_k"] else 1}, Lp={lv["fir_p"]["L"] if lv["fir_p"] else 1}')
    img = _hip.DeviceArray((grid['nx'], grid['ny']), np.complex64)
    for i in range(args.warmup):
        former(hist, out=img)
    _hip.check(_hip.lib().fsh_synchronize())
    log(f'warm: group of {former._form.info()["group"]} first-level children')
    step_s = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        t1 = time.perf_counter()
        former(hist, out=img)
        _hip.check(_hip.lib().fsh_synchronize())
        step_s.append(time.perf_counter() - t1)
    dt = (time.perf_counter() - t0) / args.steps
    pairs = grid['nx'] * grid['ny'] * col.Np
    log(f'{dt * 1e3:.1f} ms per image (median {np.median(step_s) * 1e3:.1f}), {pairs / dt / 1e9:.1f} Gpixel.pulses/s')

    # host to host (pageable history in, image out): the PCIe-inclusive figure, reported beside the device rate
    h2h = None
    if args.host_steps > 0 and not args.no_profile:
        S_host = dS.to_host() if hist_host is None else hist_host
        former(S_host)
        hs = []
        for i in range(args.host_steps):
            t0 = time.perf_counter()
            former(S_host)
            hs.append(time.perf_counter() - t0)
        h2h = float(np.median(hs))
        del S_host
        log(f'host to host: median {h2h * 1e3:.1f} ms per image of {len(hs)} ({pairs / h2h / 1e9:.1f} Gpixel.pulses/s)')

    if args.calibrate_copy:
        # the plain copy kernel on 1 GiB, three times: the byte counters' reference point
        src, dst = _hip.DeviceArray((1 << 27,), np.complex64), _hip.DeviceArray((1 << 27,), np.complex64)
        for _ in range(3):
            _hip.check(_hip.lib().fsh_copy(_hip.ptr(dst), _hip.ptr(src), 1 << 30, None))
        _hip.check(_hip.lib().fsh_synchronize())
        log('copy kernel: 3 x 1 GiB')
    if args.no_profile:
        print(json.dumps(dict(metric=METRIC, value=round(pairs / dt / 1e9, 2), ms_per_step=round(dt * 1e3, 2))), flush=True)
        return

    # per-launch device times (events on the former's stream) of one more call
    former._form.set_profile(True)
    former(hist, out=img)
    prof = former._form.profile()
    former._form.set_profile(False)
    agg = {}
    for st, lvl, ms in prof:
        a = agg.setdefault((st, lvl), [0.0, 0])
        a[0] += ms
        a[1] += 1
    flops, nbytes = stage_flops(plan), stage_bytes(plan)
    stages = {}
    for k, v in sorted(agg.items(), key=lambda kv: -kv[1][0]):
        sec = v[0] * 1e-3
        stages[f'{k[0]}[{k[1]}]'] = dict(ms=round(v[0], 3), launches=v[1],
                                         tflops=round(flops[k] / sec / 1e12, 2) if k in flops else None,
                                         gbs=round(nbytes[k] / sec / 1e9, 1) if k in nbytes else None)
    for k, v in stages.items():
        log(f'  {k}: {v}')
    # the dominant kernel: the (stage, level) with the most device time. Its roof: the float32 matrix rate where its
    # intensity (algorithmic flops over algorithmic bytes) is above that rate's ridge (157.3 TFLOP/s over 8 TB/s,
    # 19.7 flop/B) -- the path computes float32-class sums (bfloat16 x 3 on the matrix cores, float32 on VALU) --
    # else HBM. Both figures are kept.
    dom = max((k for k in agg if k in nbytes), key=lambda k: agg[k][0])
    label = f'{dom[0]}[{dom[1]}]'
    dom_ms = agg[dom][0] / agg[dom][1]
    dom_bytes = nbytes[dom] / agg[dom][1]
    dom_flops = flops.get(dom, 0.0) / agg[dom][1]
    gbs = dom_bytes / (dom_ms * 1e-3)
    tfs = dom_flops / (dom_ms * 1e-3)
    if dom_bytes > 0 and dom_flops / dom_bytes > FP32_MATRIX_PEAK / HBM_PEAK:
        roofline = dict(bound='mfma', achieved=round(tfs / 1e12, 2), peak=FP32_MATRIX_PEAK / 1e12, unit='TFLOP/s',
                        frac=round(tfs / FP32_MATRIX_PEAK, 4))
    else:
        roofline = dict(bound='hbm', achieved=round(gbs / 1e9, 1), peak=HBM_PEAK / 1e9, unit='GB/s',
                        frac=round(gbs / HBM_PEAK, 4))
    roofline.update(traffic=measured_traffic(label, former, args), kernel=label, bytes_per_launch=round(dom_bytes),
                    flop_per_launch=round(dom_flops), ms_per_launch=round(dom_ms, 4),
                    intensity_flop_per_byte=round(dom_flops / max(dom_bytes, 1), 1), hbm_gbs=round(gbs / 1e9, 1),
                    hbm_frac=round(gbs / HBM_PEAK, 4), tflops=round(tfs / 1e12, 2))

    # error against float64 exact backprojection on seeded tiles
    t0 = time.perf_counter()
    tiles, pts, side = accuracy_points(grid, args.tiles, 256, seed=7)
    nfft = 1 << int(np.ceil(np.log2(args.upsample * col.K)))
    wp, wk = _window(col.Np, col.K)
    ref_vals = bp_points_f64_device(dS, col.ant, np.linalg.norm(col.ant, axis=1), col.fmin, col.df, pts, nfft, True, wp, wk,
                                    chunk=max(1, (1 << 30) // (nfft * 16)))
    himg = img.to_host()
    got = np.concatenate([himg[i0:i0 + side, j0:j0 + side].ravel() for i0, j0 in tiles])
    err = rel_db(got, ref_vals)
    log(f'error vs float64 exact backprojection ({len(tiles)} tiles of {side}^2, cubic reads at {args.upsample}x): '
        f'{err:.2f} dB ({time.perf_counter() - t0:.1f} s)')

    # the learned form beside the plain one: the committed filters of 8 taps per phase trained on two other scenes
    # of this workload (fastsar_amd/data/learned_c2_T8.npz), same plan otherwise
    learned_form = None
    lpath = os.path.join(ROOT, 'fastsar_amd', 'data', f'learned_c2_T8.npz')
    if (not args.small and not args.no_learned and taps is None and args.T == 8 and args.levels == 4 and args.atten == 80.0
            and args.history == 'complex64' and args.precision == 'float32' and os.path.exists(lpath)):
        from fastsar_amd import learned
        lt = learned.LearnedTaps.load(lpath)
        limg = _hip.DeviceArray((grid['nx'], grid['ny']), np.complex64)

        def timed(precision):
            # warm calls, then the timed ones; -> (former, seconds per image, error against the checker in dB)
            f = ImageFormer(col.ant, col.fmin, col.df, col.K, **grid, T=args.T, levels=args.levels, atten=args.atten,
                            taps=lt, precision=precision)
            for _ in range(args.warmup):
                f(hist, out=limg)
            _hip.check(_hip.lib().fsh_synchronize())
            t1 = time.perf_counter()
            for _ in range(args.steps):
                f(hist, out=limg)
            _hip.check(_hip.lib().fsh_synchronize())
            sec = (time.perf_counter() - t1) / args.steps
            lh = limg.to_host()
            lgot = np.concatenate([lh[i0:i0 + side, j0:j0 + side].ravel() for i0, j0 in tiles])
            return f, sec, rel_db(lgot, ref_vals)

        lf, ldt, lerr = timed('float32')
        lplan = lf._plan
        del lf
        learned_form = dict(ms_per_step=round(ldt * 1e3, 2), value=round(pairs / ldt / 1e9, 2),
                            error_db=round(lerr, 2), speedup=round(dt / ldt, 3),
                            range_taps=[l['fir_k']['L'] if l['fir_k'] else 1 for l in lplan['levels']],
                            pulse_taps=[l['fir_p']['L'] if l['fir_p'] else 1 for l in lplan['levels']],
                            kaiser_kept=sorted(f'{a}[{b}]' for b, a in lt.keep))
        # the same filters with the two-pass levels' rows stored in float16
        hf, hdt, herr = timed('float16')
        del hf
        learned_form['float16'] = dict(ms_per_step=round(hdt * 1e3, 2), value=round(pairs / hdt / 1e9, 2),
                                       error_db=round(herr, 2), speedup=round(dt / hdt, 3))
        log(f'learned form: {learned_form}')
        del limg

    # the same image with every stage on float32 VALU kernels (no matrix cores, no bfloat16 operands): the plain float32
    # form that the matrix-core form (bfloat16 x 3 operands, float32 accumulation) and the learned float16 form stand
    # beside, timed on one warm and two timed calls
    valu_form = None
    if not args.small and not args.no_valu_form and taps is None and args.precision == 'float32' and args.history == 'complex64':
        keep = {k: os.environ.get(k) for k in ('FSH_ROT_KERNEL', 'FSH_LEVEL_KERNEL')}
        os.environ['FSH_ROT_KERNEL'], os.environ['FSH_LEVEL_KERNEL'] = 'valu', 'split'
        try:
            vimg = _hip.DeviceArray((grid['nx'], grid['ny']), np.complex64)
            vf = ImageFormer(col.ant, col.fmin, col.df, col.K, **grid, T=args.T, levels=args.levels, atten=args.atten)
            vf(hist, out=vimg)
            _hip.check(_hip.lib().fsh_synchronize())
            t1 = time.perf_counter()
            for _ in range(2):
                vf(hist, out=vimg)
            _hip.check(_hip.lib().fsh_synchronize())
            vdt = (time.perf_counter() - t1) / 2
            vh = vimg.to_host()
            verr = rel_db(np.concatenate([vh[i0:i0 + side, j0:j0 + side].ravel() for i0, j0 in tiles]), ref_vals)
            del vf, vimg, vh
        finally:
            for k, v in keep.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
        valu_form = dict(ms_per_step=round(vdt * 1e3, 2), value=round(pairs / vdt / 1e9, 2), error_db=round(verr, 2),
                         matrix_core_speedup=round(vdt / dt, 3),
                         learned_float16_speedup=(round(vdt / (learned_form['float16']['ms_per_step'] * 1e-3), 3)
                                                  if learned_form else None),
                         kernels='rot_fir_k VALU (float32 taps per row), fir_p streaming, level_fused VALU, final stage VALU')
        log(f'float32 VALU form: {valu_form}')

    exact = None
    if not args.no_exact and not args.small and args.history == 'complex64':
        exact = exact_leg(col, grid, dS, tiles, side, ref_vals, args.steps)

    batch = None
    if not args.no_batch and not args.small and args.history == 'complex64':
        # the 48-image batch in its bounded form (no candidate timing); bench.py --batch autotunes it. The image's
        # former and its buffers are released first.
        del himg, img, former
        bl = run_batch(args, tune=False)
        batch = dict(bl['config'], ms=bl['ms_per_step'], value=bl['value'], metric=bl['metric'])

    cpu = None
    if not args.no_cpu:
        try:
            import oracle
            if oracle.have_ref():
                t0 = time.perf_counter()
                cpu = cpu_baseline(col, grid, dS.to_host(), args.cpu_images, tiles, side, ref_vals)
                log(f'cpu baseline: {cpu} ({time.perf_counter() - t0:.1f} s)')
        except Exception as e:          # the baseline is reported, not required
            log(f'cpu baseline skipped: {type(e).__name__}: {e}')

    out = dict(metric=METRIC, value=round(pairs / dt / 1e9, 2), unit='Gpixel·pulses/s', n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=round(dt * 1e3, 2), higher_is_better=True, scaling='weak',
               vs_baseline=None, dtype='f32', data='synthetic',
               config=dict(workload=(f'{"512x512 from 512x1024" if args.small else "12288x8192 image from 16384 pulses x 8192 samples"}'
                                     f' {args.history}, factorized backprojection, T={form_T}, {form_levels} levels, '
                                     + (f'learned filters of {taps.T} taps per phase' if taps else f'{args.atten:g} dB filters')),
                           history=args.history, precision=args.precision,
                           s_per_image=round(dt, 4), device_median_ms=round(float(np.median(step_s)) * 1e3, 2),
                           error_db=round(err, 2), error_tiles=f'{len(tiles)} x {side}^2',
                           predicted_error_db=round(form_pred, 2), parallelism='single GPU',
                           host_to_host_ms=None if h2h is None else round(h2h * 1e3, 1),
                           host_to_host_steps=args.host_steps if h2h is not None else 0, learned_form=learned_form,
                           fp32_valu_form=valu_form,
                           exact_form=exact, batch_form=batch),
               roofline=roofline, cpu_baseline=cpu, stages=stages)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
