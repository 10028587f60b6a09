"""The range gate: a wide-band phase history cut, pulse by pulse, to the range swath an image spans, before
formation (fastsar/api.py:272-300, 525-564; the mosaic's patch gate, fastsar/patches.py).

Each pulse's spectrum (K samples spaced df) is padded by a guard of an eighth of the band on either side to Kp
samples and transformed to range, where bin n lies n c / (2 Kp df) beyond the pulse's reference range. Only the N
bins the image's points reach over the aperture (with a margin of six resolution cells of the guard band) are
kept, the first N - N/2 and the last N/2, and transformed back: the same pulse on N frequencies spaced Kp df / N,
of which those over the band and its guards, tapered to zero across the guards, are kept. A gated history of
kept = k1 - k0 samples forms the same image as the whole history, with a first level that decimates far less.

    g = gate_plan(ant, ref, fmin, df, K, nx, ny, spx, spy, e1, e2)
    gate = RangeGate(g, P, K, window=(wp, wk))          # device buffers and transform plans, once
    Sg = gate(S)                                        # DeviceArray [P, kept] complex64, windowed
"""
import ctypes

import numpy as np

from . import _hip
from .plan import C


def taper(f, band, guard):
    """The gated spectrum's weight at frequencies f: 1 on the band [band[0], band[1]], falling to 0 over `guard` on
    either side along half a period of a cosine, 0 beyond (the mosaic's taper, fastsar/patches.py:156-159)."""
    beyond = np.maximum(np.maximum(band[0] - f, f - band[1]), 0.0) / guard      # guards crossed, 0 on the band
    return 0.5 * (1 + np.cos(np.pi * np.minimum(beyond, 1.0)))


def range_reach(ant, ref, nx, ny, spx, spy, e1, e2, pulses=257, lattice=5):
    """The largest |range - reference range| from the image's points to the antenna: the image sampled on a
    lattice x lattice grid of points spanning it, the track at `pulses` positions spread evenly over it."""
    sel = np.unique(np.linspace(0, len(ant) - 1, min(len(ant), pulses)).astype(np.int64))
    track = ant[sel]
    r_ref = np.linalg.norm(track, axis=1) if ref is None else np.asarray(ref, np.float64)[sel]
    frac = np.linspace(-0.5, 0.5, lattice)
    pts = ((frac * nx * spx)[:, None, None] * e1 + (frac * ny * spy)[None, :, None] * e2).reshape(-1, 3)
    return np.abs(np.linalg.norm(pts[None] - track[:, None], axis=2) - r_ref[:, None]).max()


def gate_plan(ant, ref, fmin, df, K, nx, ny, spx, spy, e1, e2):
    """The gate of an nx x ny image about the origin. -> dict(npl: samples padded below the band, Kp: padded
    length, N: range bins kept, k0, k1: the gated spectrum's samples kept, taper [k1 - k0] float32 (with the scale
    Kp / N, which keeps the image's amplitude: the gated history sums N / Kp times fewer samples), fmin, df: of the
    gated history). Decisions as fastsar/api.py:272-300 makes them.

    The band's guard is an eighth of its width on either side, npl = ceil(guard / df) samples, padded to Kp, a fast
    transform length. A range bin is c / (2 Kp df); the kept bins cover the image's range reach (range_reach) plus six
    resolution cells of the guard band, on both sides of zero delay: N, even and of a fast length. Transformed back,
    the N frequencies are spaced Kp df / N from the padded band's first, and those the taper does not zero are kept."""
    from scipy.fft import next_fast_len
    ant = np.asarray(ant, np.float64)
    e1, e2 = np.asarray(e1, np.float64), np.asarray(e2, np.float64)
    lo, hi = fmin, fmin + (K - 1) * df
    guard = 0.125 * (hi - lo)
    npl = int(np.ceil(guard / df))
    Kp = next_fast_len(K + 2 * npl)
    reach = range_reach(ant, ref, nx, ny, spx, spy, e1, e2) + 6 * C / (2 * guard)
    N = next_fast_len(2 * int(np.ceil(reach * 2 * Kp * df / C)))
    N += N & 1
    spacing, first = Kp * df / N, fmin - npl * df
    weight = taper(first + spacing * np.arange(N), (lo, hi), guard)
    live = np.flatnonzero(weight > 0)
    k0, k1 = int(live[0]), int(live[-1]) + 1
    return dict(npl=npl, Kp=int(Kp), N=int(N), k0=k0, k1=k1, taper=(weight[k0:k1] * (Kp / N)).astype(np.float32),
                fmin=float(first + k0 * spacing), df=float(spacing))


class GateDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ('P', 'K', 'npl', 'Kp', 'N', 'k0', 'k1')] + \
               [('taper', _hip.P_f32), ('wp', _hip.P_f32), ('wk', _hip.P_f32)]


def _bind():
    L = _hip.lib()
    if not getattr(L, '_gate_bound', False):
        vp, ci = ctypes.c_void_p, ctypes.c_int
        L.fsh_gate_create.argtypes = [ctypes.POINTER(GateDesc), ctypes.POINTER(vp)]
        L.fsh_gate_run.argtypes = [vp, vp, ci, vp, _hip.P_i32]
        L.fsh_gate_destroy.argtypes = [vp]
        L.fsh_gate_set_ramp.argtypes = [vp, vp]
        for n in ('fsh_gate_create', 'fsh_gate_run', 'fsh_gate_destroy', 'fsh_gate_set_ramp'):
            getattr(L, n).restype = ci
        L._gate_bound = True
    return L


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, np.float32)


class RangeGate:
    """The gate of a plan (gate_plan) on the device for histories of P pulses and K samples: window, pad, inverse
    transform, range cut, forward transform and taper, in chunks of pulses, on the library's own stream (hipFFT
    transforms). gate(S) takes a host array or a DeviceArray [P, K] complex64 and returns a DeviceArray
    [P, k1 - k0]; a NaN or inf in S raises ValueError naming its pulse."""

    def __init__(self, g, P, K, window=None):
        L = _bind()
        self.g, self.P, self.K = g, int(P), int(K)
        self.kept = int(g['k1'] - g['k0'])
        tap = _f32(g['taper'])
        wp, wk = (None, None) if window is None else (_f32(window[0]), _f32(window[1]))
        if window is not None and (wp.shape != (self.P,) or wk.shape != (self.K,)):
            raise ValueError(f'window must be ([{self.P}], [{self.K}]), got {wp.shape}, {wk.shape}')
        d = GateDesc(P=self.P, K=self.K, npl=int(g['npl']), Kp=int(g['Kp']), N=int(g['N']), k0=int(g['k0']), k1=int(g['k1']))
        d.taper = tap.ctypes.data_as(_hip.P_f32)
        if wp is not None:
            d.wp, d.wk = wp.ctypes.data_as(_hip.P_f32), wk.ctypes.data_as(_hip.P_f32)
        h = ctypes.c_void_p()
        _hip.check(L.fsh_gate_create(ctypes.byref(d), ctypes.byref(h)), 'range gate')
        self._h = h.value

    def set_ramp(self, ab):
        """Per pulse (a_p, b_p) in cycles [P, 2] (or None): sample k of pulse p times exp(2 pi i (a_p + b_p k)) before
        the window and the gate (a re-reference of the pulses' ranges)."""
        L = _bind()
        if ab is None:
            _hip.check(L.fsh_gate_set_ramp(ctypes.c_void_p(self._h), None), 'range gate ramp')
            return
        ab = np.ascontiguousarray(ab, np.float64)
        if ab.shape != (self.P, 2) or not np.all(np.isfinite(ab)):
            raise ValueError(f'the ramp must be [{self.P}, 2] finite cycles, got shape {ab.shape}')
        _hip.check(L.fsh_gate_set_ramp(ctypes.c_void_p(self._h), ab.ctypes.data_as(ctypes.c_void_p)), 'range gate ramp')

    def __call__(self, S, out=None):
        L = _bind()
        on_dev = isinstance(S, _hip.DeviceArray)
        if on_dev:
            if S.shape != (self.P, self.K) or S.dtype != np.complex64:
                raise ValueError(f'phase history must be {(self.P, self.K)} complex64, got {S.shape} {S.dtype}')
        else:
            S = np.ascontiguousarray(S, np.complex64)
            if S.shape != (self.P, self.K):
                raise ValueError(f'phase history must be {(self.P, self.K)}, got {S.shape}')
        if out is None:
            out = _hip.DeviceArray((self.P, self.kept), np.complex64)
        elif not isinstance(out, _hip.DeviceArray) or out.shape != (self.P, self.kept) or out.dtype != np.complex64:
            raise ValueError(f'out must be a DeviceArray {(self.P, self.kept)} complex64')
        bad = ctypes.c_int(-1)
        _hip.check(L.fsh_gate_run(ctypes.c_void_p(self._h), _hip.ptr(S), int(on_dev), _hip.ptr(out), ctypes.byref(bad)),
                   'range gate')
        return out

    def __del__(self):
        try:
            if getattr(self, '_h', None) and _hip._lib is not None:
                _hip._lib.fsh_gate_destroy(ctypes.c_void_p(self._h))
        except Exception:
            pass
        self._h = None
