"""Many images of uneven size in one launch, and the autotuner that picks each image's final tile and levels.

    formers = [autotune(col.ant, col.fmin, col.df, col.K, **grid, S=dS) for ...]    # one per image
    b = Batch(formers, histories)                   # device histories; images allocated here
    r = b.run()                                     # one hipGraph launch: r['ms'], r['done_ms'], r['bad']
    img = b.images[i].to_host()

The batch captures every former's launches (window and NaN scan, the levels in groups, the final stage) into one
hipGraph, spread over forked streams so that small images run beside each other (by default balanced: the images,
timed one call each, each given in turn, largest first, to the stream with the least time so far); replays launch
the whole batch at once. No reference counterpart: the reference forms its patches one former call at a time
(fastsar/patches.py:790-870), which is what `run_loop` does here for comparison.
"""
import ctypes
import time

import numpy as np

from . import _hip

CANDIDATES = ((8, 4), (8, 3), (16, 3), (16, 4), (4, 4), (32, 3))     # (final tile T, levels)
_TUNED = {}                                                         # (P, K, nx, ny) -> candidates, fastest first


def _bind():
    L = _hip.lib()
    if not getattr(L, '_batch_bound', False):
        vp, ci = ctypes.c_void_p, ctypes.c_int
        L.fsh_ffbp_enqueue.argtypes = [vp, vp, vp, vp]
        L.fsh_ffbp_bad_pulse.argtypes = [vp, _hip.P_i32]
        L.fsh_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), ci, ci,
                                       ctypes.POINTER(ci), ctypes.POINTER(vp)]
        L.fsh_batch_run.argtypes = [vp, _hip.P_f32, _hip.P_f32, _hip.P_i32]
        L.fsh_batch_destroy.argtypes = [vp]
        for n in ('fsh_ffbp_enqueue', 'fsh_ffbp_bad_pulse', 'fsh_batch_create', 'fsh_batch_run', 'fsh_batch_destroy'):
            getattr(L, n).restype = ci
        L._batch_bound = True
    return L


def autotune(ant, fmin, df, K, nx, ny, spx, spy, e1=(1.0, 0.0, 0.0), e2=(0.0, 1.0, 0.0), *, S, target_db=-55.0,
             atten=80.0, candidates=CANDIDATES, window=True, trials=2):
    """The fastest ImageFormer among `candidates` (final tile T, levels) whose predicted final-stage error meets
    target_db, timed on the device history S (a DeviceArray [P, K]) after one warm call. The ranking is kept per
    (P, K, nx, ny): later images of the same shape build only the first ranked candidate that meets their own
    target. -> the former (warm), with former.tuned = dict(T, levels, ms, tried)."""
    from .api import ImageFormer
    key = (len(ant), int(K), int(nx), int(ny))
    img = _hip.DeviceArray((nx, ny), np.complex64)
    ranked = _TUNED.get(key)
    if ranked is not None:
        for T, lv in ranked:
            f = ImageFormer(ant, fmin, df, K, nx, ny, spx, spy, e1, e2, T=T, levels=lv, atten=atten, window=window)
            if f.predicted_error_db <= target_db:
                f(S, out=img)
                f.tuned = dict(T=T, levels=lv, ms=None, tried=0)
                return f
    timed = []
    best = None
    for T, lv in candidates:
        try:
            f = ImageFormer(ant, fmin, df, K, nx, ny, spx, spy, e1, e2, T=T, levels=lv, atten=atten, window=window)
        except ValueError:                  # a plan that cannot split the grid this way
            continue
        if f.predicted_error_db > target_db:
            continue
        f(S, out=img)
        _hip.check(_hip.lib().fsh_synchronize())
        t0 = time.perf_counter()
        for _ in range(trials):
            f(S, out=img)
        ms = (time.perf_counter() - t0) / trials * 1e3
        timed.append((ms, T, lv))
        if best is None or ms < best[0]:
            best = (ms, f)
        else:
            del f
    if best is None:
        raise ValueError(f'no candidate (T, levels) meets the error target {target_db} dB for this image')
    timed.sort()
    _TUNED[key] = [(T, lv) for _, T, lv in timed]
    f = best[1]
    f.tuned = dict(T=f.T, levels=f.levels, ms=best[0], tried=len(timed))
    return f


def stream_order(cost, ns):
    """Images with the given costs over ns streams, largest first, each to the stream with the least cost so far
    (the lowest-numbered on ties). -> (order, stream): the capture order (order[p] = the image captured p-th) and
    the stream of each captured position."""
    n = len(cost)
    order = sorted(range(n), key=lambda i: (-cost[i], i))
    load, stream = [0.0] * ns, []
    for i in order:
        s = min(range(ns), key=lambda k: (load[k], k))
        load[s] += cost[i]
        stream.append(s)
    return order, stream


class Batch:
    """One hipGraph over the formers' images: formers (warm ImageFormers without a range gate), histories (device
    complex64 [P, K] each), images (device [nx, ny] each, allocated when None); streams: forked streams the images
    are spread over; balance: give the images to the streams by their single-call times (largest first, each to the
    stream with the least time so far) instead of round-robin in the given order. Replayable: each run() forms every
    image again from its history as it then is."""

    def __init__(self, formers, histories, images=None, streams=4, balance=True):
        L = _bind()
        n = len(formers)
        if n < 1 or len(histories) != n:
            raise ValueError('a batch needs one history per former')
        for f, S in zip(formers, histories):
            if getattr(f, '_gate', None) is not None:
                raise ValueError('a gated former is formed by its own call, not in a batch')
            if not isinstance(S, _hip.DeviceArray) or S.shape != (f.P, f.K) or S.dtype != np.complex64:
                raise ValueError('batch histories must be complex64 DeviceArrays of each former\'s [P, K]')
            if f._form.info()['group'] < 1:
                f(S)                                               # the group buffers are made by a first call
        self.formers, self.histories = list(formers), list(histories)
        if images is not None:
            if len(images) != n:
                raise ValueError('a batch needs one image per former')
            for f, im in zip(formers, images):
                if not isinstance(im, _hip.DeviceArray):
                    raise ValueError('batch images must be DeviceArrays (the graph writes them on the device)')
                _hip.output(im, (f._form.nx, f._form.ny))
        self.images = list(images) if images is not None else [_hip.DeviceArray((f._form.nx, f._form.ny), np.complex64)
                                                               for f in formers]
        # the capture order (image order[p] captured p-th) and each position's stream
        ns = max(1, min(int(streams), n))
        self.order, stream = list(range(n)), None
        if balance and n > ns:
            cost = []                      # each image's launches timed alone (enqueued: a bad history raises nothing)
            for f, S, im in zip(self.formers, self.histories, self.images):
                _hip.check(L.fsh_synchronize())
                t0 = time.perf_counter()
                _hip.check(L.fsh_ffbp_enqueue(f._form.h, S.ptr, im.ptr, None), 'ffbp_enqueue')
                _hip.check(L.fsh_synchronize())
                cost.append(time.perf_counter() - t0)
            self.order, stream = stream_order(cost, ns)
        vp = ctypes.c_void_p
        hs = (vp * n)(*[self.formers[i]._form.h for i in self.order])
        ss = (vp * n)(*[self.histories[i].ptr for i in self.order])
        ims = (vp * n)(*[self.images[i].ptr for i in self.order])
        h = vp()
        so = None if stream is None else (ctypes.c_int * n)(*stream)
        _hip.check(L.fsh_batch_create(hs, ss, ims, n, int(streams), so, ctypes.byref(h)), 'batch_create')
        self._h = h.value
        self.n = n

    def run(self):
        """One launch of the batch, waited for -> dict(ms: host clock for the batch, done_ms [n]: each image's
        completion after the batch's start by the device clock, bad [n]: first non-finite pulse or -1)."""
        L = _bind()
        ms = ctypes.c_float()
        done = (ctypes.c_float * self.n)()
        bad = (ctypes.c_int * self.n)()
        _hip.check(L.fsh_batch_run(ctypes.c_void_p(self._h), ctypes.byref(ms), done, bad), 'batch_run')
        d, b = np.empty(self.n), [0] * self.n
        for p, i in enumerate(self.order):                         # back to the caller's order
            d[i], b[i] = done[p], bad[p]
        return dict(ms=ms.value, done_ms=d, bad=b)

    def run_loop(self):
        """The same images formed by one former call each, in order (the reference's way) -> dict(ms, each_ms)."""
        each = []
        t0 = time.perf_counter()
        for f, S, im in zip(self.formers, self.histories, self.images):
            t1 = time.perf_counter()
            f(S, out=im)
            each.append((time.perf_counter() - t1) * 1e3)
        return dict(ms=(time.perf_counter() - t0) * 1e3, each_ms=np.array(each))

    def __del__(self):
        try:
            if getattr(self, '_h', None) and _hip._lib is not None:
                _hip._lib.fsh_batch_destroy(ctypes.c_void_p(self._h))
        except Exception:
            pass
        self._h = None
