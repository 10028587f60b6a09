"""The former API of FastSAR's image formation, on the 'hip' backend (fastsar/api.py).

    import fastsar_amd as fastsar
    img = fastsar.form_image(S, ant, fmin, df, nx, ny, spx, spy, e1, e2)          # factorized backprojection
    img = fastsar.form_image(..., algorithm='bp')                                 # exact backprojection
    former = fastsar.ImageFormer(ant, fmin, df, K, nx, ny, spx, spy, e1, e2); img = former(S)

S is the phase history [pulses, samples] (complex, frequency domain, motion compensated to the scene reference
point), ant the antenna phase centres [pulses, 3] in a frame whose origin is that point, fmin and df the first
frequency and the sample spacing (Hz), and nx, ny, spx, spy, e1, e2 the output grid: pixel (i, j) lies at
(i - nx/2) spx e1 + (j - ny/2) spy e2. The result is a complex64 image [nx, ny], the plain sum over pulses and
samples of the windowed history times exp(+j 4 pi f dR / c).

Backends: 'hip' (hand-written HIP kernels on an AMD Instinct GPU, gfx950) and 'auto' (the same). The reference's
other backend names are recognised and refused with the reason.
"""
import ctypes
import math
import warnings

import numpy as np

from . import gate as _gate
from . import memory as _mem
from . import plan as _plan
from .sim import Collect

C = _plan.C
BACKENDS = ('auto', 'hip', 'tpu', 'cuda', 'cpu', 'jax', 'metal')
_OTHER = {'tpu': 'JAX on a Cloud TPU', 'cuda': 'CuPy on an Nvidia GPU', 'cpu': 'the C++/OpenMP kernels',
          'jax': 'the JAX program', 'metal': 'Apple GPUs'}


def _window(P, K, sll=35.0, nbar=4):
    """Taylor windows (-35 dB, nbar 4, unnormalized) along pulses and samples, float32. (fastsar/api.py:54-56)"""
    from scipy.signal.windows import taylor
    return (taylor(P, nbar=nbar, sll=sll, norm=False).astype(np.float32),
            taylor(K, nbar=nbar, sll=sll, norm=False).astype(np.float32))


def available_backends():
    """The backends this machine can run, in the order 'auto' tries them. (fastsar/api.py:66-90)"""
    from . import _hip
    return ['hip'] if _hip.available() else []


def _backend(backend):
    """backend checked against BACKENDS and this machine; 'auto' resolved. (fastsar/api.py:93-129)"""
    if backend not in BACKENDS:
        raise ValueError(f'unknown backend {backend!r}: one of {", ".join(map(repr, BACKENDS))}')
    if backend in _OTHER:
        raise ValueError(f"backend {backend!r} ({_OTHER[backend]}) belongs to the reference package; this package "
                         "forms images with backend='hip'")
    from . import _hip
    _hip.lib()                                     # raises when the library is not built
    if not _hip.available():
        raise ValueError("backend 'hip' needs an AMD GPU with ROCm; none answered")
    return 'hip'


def _check_history(S, P=None, name='phase history', finite=True):
    """S [P, K] complex with at least 2 pulses and 2 samples (finite: scanned for NaN and inf here; the hip former
    scans on the device). -> S. (fastsar/api.py:132-152)"""
    if not hasattr(S, 'ndim'):
        S = np.asarray(S)
    if S.ndim != 2:
        raise ValueError(f'{name} must be 2-D [pulses, samples], got shape {tuple(S.shape)}')
    if S.dtype.kind != 'c':
        raise TypeError(f'{name} must be complex (complex64 or complex128), got {S.dtype}')
    if P is not None and S.shape[0] != P:
        raise ValueError(f'{name} has {S.shape[0]} pulses but there are {P} antenna positions')
    if S.shape[0] < 2 or S.shape[1] < 2:
        raise ValueError(f'{name} needs at least 2 pulses and 2 samples, got shape {tuple(S.shape)}')
    for i in range(0, S.shape[0] if finite else 0, 4096):
        if not np.isfinite(complex(S[i:i + 4096].sum())):
            first = i + int(np.argmax((~np.isfinite(S[i:i + 4096])).any(1)))
            raise ValueError(f'{name} has non-finite samples (NaN or inf), the first in pulse {first}; '
                             'zero them (S[~np.isfinite(S)] = 0) or drop the pulses')
    return S


def _check_positions(a, name='antenna positions', P=None):
    """a [P, 3] finite float64 with at least 2 rows. (fastsar/api.py:155-166)"""
    a = np.asarray(a, np.float64)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError(f'{name} must be [pulses, 3], got shape {a.shape}')
    if P is not None and len(a) != P:
        raise ValueError(f'{name}: {len(a)} rows for {P} pulses')
    if len(a) < 2:
        raise ValueError(f'{name}: at least 2 pulses are needed, got {len(a)}')
    if not np.isfinite(a).all():
        raise ValueError(f'{name}: non-finite values (NaN or inf)')
    return a


def _check_grid(nx, ny, spx, spy, e1, e2):
    """Pixel counts, spacings and orthonormal axes -> nx, ny (int), e1, e2 (float64). (fastsar/api.py:169-183)"""
    for n, v in (('nx', nx), ('ny', ny)):
        if np.ndim(v) != 0 or not np.isfinite(v) or v < 1 or int(v) != v:
            raise ValueError(f'{n} must be a positive integer, got {v!r}')
    for n, v in (('spx', spx), ('spy', spy)):
        if not np.isfinite(v) or v <= 0:
            raise ValueError(f'{n} must be a positive spacing in meters, got {v!r}')
    e1, e2 = np.asarray(e1, np.float64), np.asarray(e2, np.float64)
    if e1.shape != (3,) or e2.shape != (3,):
        raise ValueError(f'e1 and e2 must be 3-vectors, got shapes {e1.shape} and {e2.shape}')
    if abs(e1 @ e1 - 1) > 1e-6 or abs(e2 @ e2 - 1) > 1e-6 or abs(e1 @ e2) > 1e-6:
        raise ValueError(f'e1 and e2 must be orthonormal (unit length, perpendicular), got {e1.tolist()} and {e2.tolist()}')
    return int(nx), int(ny), e1, e2


def final_weights(plan, weight, P, grad=False, points=None, nodes=3, first=0):
    """Weight of each final subaperture at each final tile [ntiles, Pf] (float32): weight(points [m, 3], pulses [n]
    int) -> [n, m] per-pulse weights, averaged over the pulses each final subaperture spans (its centre composed
    through the levels' pulse grids, width the product of the decimations) by 3-point Gauss-Legendre quadrature
    (nodes=1: the centre pulse). grad=True: [3, ntiles, Pf] with the gradient along e1 and e2 at the tile centres
    by central differences half a tile away. first: the collection's first pulse within the plan's pulses.
    (fastsar/api.py:203-244)"""
    if nodes not in (1, 3):
        raise ValueError(f'final_weights: nodes must be 1 or 3, got {nodes!r}')
    cen = np.asarray(plan['final']['cen'], np.float64)
    if grad:
        h1, h2 = 0.5 * plan['T'] * plan['spx'], 0.5 * plan['T'] * plan['spy']
        e1, e2 = np.asarray(plan['e1'], np.float64), np.asarray(plan['e2'], np.float64)
        w0 = final_weights(plan, weight, P, points=cen, first=first)
        m = len(cen)
        quad = np.concatenate([cen + h1 * e1, cen - h1 * e1, cen + h2 * e2, cen - h2 * e2])
        wq = final_weights(plan, weight, P, points=quad, nodes=1, first=first)
        return np.stack([w0, (wq[:m] - wq[m:2 * m]) / (2 * h1), (wq[2 * m:3 * m] - wq[3 * m:]) / (2 * h2)]).astype(np.float32)
    q = cen if points is None else points
    idx, D = np.arange(plan['final']['P'], dtype=np.float64), 1.0
    for lv in reversed(plan['levels']):
        idx = lv['Dp'] * idx + float(lv['pidx'][0])
        D *= lv['Dp']
    lo = np.clip(idx - D / 2, first - 0.5, first + P - 0.5)
    hi = np.clip(idx + D / 2, first - 0.5, first + P - 0.5)
    if nodes == 1:
        t, gw = np.array([0.0]), np.array([1.0])
    else:
        t, gw = np.array([-math.sqrt(0.6), 0.0, math.sqrt(0.6)]), np.array([5.0, 8.0, 5.0]) / 18.0
    nd = np.clip(np.rint(0.5 * (lo + hi)[:, None] + 0.5 * (hi - lo)[:, None] * t[None, :]), first, first + P - 1).astype(np.int64)
    U, inv = np.unique(nd.ravel(), return_inverse=True)
    Wt = np.asarray(weight(q, U), np.float64)[inv].reshape(len(idx), len(t), -1)
    m = gw[0] * Wt[:, 0]
    for k in range(1, len(t)):
        m = m + gw[k] * Wt[:, k]
    return np.ascontiguousarray(m.T, np.float32)


class ImageFormer:
    """Factorized backprojection set up once for a collection geometry and output grid, then called on phase
    histories: former = ImageFormer(ant, fmin, df, K, nx, ny, spx, spy, e1, e2); img = former(S).
    (fastsar/api.py:312-677)

    T: 'auto' (the largest of 32, 16, 8 and 4 whose predicted error meets target_db) or 4, 8, 16, 32. levels: None
    (three, more where three cannot split an axis) or a count. aperture_weight(points, pulses) -> [n, m]: per-pixel
    pulse weights applied in the final stage. ref [P]: each pulse's reference range when it is not |ant|.
    precision: 'float32', or 'float16' (the two-pass levels' intermediate rows stored in float16). range_gate: True
    cuts each pulse to the range bins the image spans before formation (a
    wide-band history for a narrow image: fewer samples, a first level that decimates less); 'auto' and False do
    not gate, since the hip first level holds any range filter.
    atten: stopband attenuation (dB) of the levels' Kaiser decimation filters; 70 (the reference's design, whose
    plans this reproduces bit for bit) sets an image error floor near -60 dB, 80 lowers it by about 10 dB for
    filters about 16 percent longer. taps: learned decimation filters (fastsar_amd.learned.LearnedTaps, trained for
    the plan's number of levels) in place of the Kaiser ones."""

    def __init__(self, ant, fmin, df, K, nx, ny, spx, spy, e1=(1.0, 0.0, 0.0), e2=(0.0, 1.0, 0.0), *, backend='auto',
                 precision='float32', window=True, T='auto', levels=None, pmax=0.4, target_db=-40.0, aperture_weight=None,
                 ref=None, range_gate='auto', padded=None, atten=70.0, taps=None):
        if not (np.isfinite(atten) and 40.0 <= float(atten) <= 140.0):
            raise ValueError(f'atten must be a stopband attenuation between 40 and 140 dB, got {atten!r}')
        self.ant = _check_positions(ant)
        self.P, self.K = self.ant.shape[0], int(K)
        if self.K < 2:
            raise ValueError(f'K must be at least 2 frequency samples, got {K}')
        if not (np.isfinite(fmin) and np.isfinite(df) and fmin > 0 and df > 0):
            raise ValueError(f'fmin and df must be positive frequencies in Hz, got {fmin!r} and {df!r}')
        self.window = window
        self._fmin, self._df = float(fmin), float(df)
        col = Collect(fmin=float(fmin), df=float(df), K=self.K, ant=self.ant, res=0.5)
        nx, ny, e1, e2 = _check_grid(nx, ny, spx, spy, e1, e2)
        self.backend = _backend(backend)
        if precision not in ('float32', 'float16'):
            raise ValueError(f"hip precision: 'float32' or 'float16', got {precision!r}")
        self.precision = precision
        self.predicted_error_db = None
        if T == 'auto':
            # the largest of 32, 16, 8, 4 whose predicted error meets target_db: the hip final stage takes all four
            # (the reference's cpu and cuda stages take 32 and 16, its jax program 8 too; fastsar/ffbp2.py:300-307)
            fmax = float(fmin) + self.K * float(df)
            T, err = _plan.choose_T(self.ant, fmax, nx, ny, spx, spy, e1, e2, target_db, choices=(32, 16, 8, 4))
            if err > target_db:
                warnings.warn(f'predicted error {err:.1f} dB at T={T} misses target {target_db:.1f} dB '
                              '(short range for this pixel size)')
            self.predicted_error_db = float(err)
        elif T not in (4, 8, 16, 32):
            raise ValueError(f'T must be 4, 8, 16, 32 or "auto", got {T!r}')
        else:
            fmax = float(fmin) + self.K * float(df)
            self.predicted_error_db = float(_plan.final_phase_error(self.ant, fmax, nx, ny, spx, spy, e1, e2, T))
        self.T = T
        nlev = 3 if levels is None else int(levels)
        if ref is not None and np.shape(ref) != (self.P,):
            raise ValueError(f'ref must hold one range per pulse ({self.P}), got shape {np.shape(ref)}')
        if ref is not None and not np.all(np.isfinite(ref)):
            raise ValueError('ref has non-finite values')
        if range_gate not in ('auto', True, False):
            raise ValueError(f"range_gate must be 'auto', True or False, got {range_gate!r}")
        if padded is not None and (aperture_weight is None or levels is None or len(padded[2]) != levels
                                   or not 0 <= padded[0] <= self.P - padded[1]):
            raise ValueError('padded=(first, count, pulse) goes with an aperture_weight, the levels of the plan that '
                             f'pulse is from and an aperture within the {self.P} pulses, got {padded!r}')
        def planned(col, nlev):
            while True:
                try:
                    return _plan.make_plan(col, nx, ny, spx, spy, T=T, nlev=nlev, pmax=pmax, atten=float(atten), e1=e1,
                                           e2=e2, pulse=padded and padded[2]), nlev
                except ValueError as e:
                    if levels is not None or 'cannot split' not in str(e) or nlev >= 8:
                        raise
                    nlev += 1
        plan, nlev = planned(col, nlev)
        self.range_gate = None
        self._gate = self._gated = None
        if range_gate is True:
            # 'auto' gates only where a backend's first level cannot hold the range filter, which the hip first
            # level always can; True gates wherever the image spans less than the swath (fastsar/api.py:404-417)
            g = _gate.gate_plan(self.ant, ref, float(fmin), float(df), self.K, nx, ny, spx, spy, e1, e2)
            kept = g['k1'] - g['k0']
            if g['N'] < g['Kp'] and kept < self.K // 2:
                gcol = Collect(fmin=g['fmin'], df=g['df'], K=kept, ant=self.ant, res=0.5)
                plan, nlev = planned(gcol, 3 if levels is None else int(levels))
                self.range_gate = dict(g, K=self.K, kept=kept, note=(
                    f'range gate: each pulse cut to the {g["N"]} range bins about the image before formation, '
                    f'{kept} samples for {self.K}'))
        if taps is not None:
            # learned decimation filters (fastsar_amd.learned.LearnedTaps) in place of the plan's Kaiser ones
            plan = taps.apply(plan)
        self._Kf = plan['levels'][0]['K']
        self.levels = nlev
        self._plan = plan
        coll = _plan.collection_arrays(plan, self.ant, ref)
        wspan = dict(P=self.P) if padded is None else dict(P=int(padded[1]), first=int(padded[0]))
        wf = None if aperture_weight is None else final_weights(plan, aperture_weight, grad=True, **wspan)
        self._weighted = wf is not None
        if window:
            self.wp, self.wk = _window(self.P, self.K)
        win = (self.wp, self.wk) if window else None
        from .ffbp_hip import make_ffbp_hip
        if self.range_gate is not None:
            # the gate windows the history; the former takes the gated one as it is
            self._gate = _gate.RangeGate(self.range_gate, self.P, self.K, window=win)
            self._form = make_ffbp_hip(plan, coll, wf=wf, window=None)
        else:
            self._form = make_ffbp_hip(plan, coll, wf=wf, window=win)
        if precision == 'float16':
            # the two-pass levels' range-filtered rows stored in float16 (scaled per call, no overflow): half their
            # traffic for a rounding near -74 dB (the reference's cuda 'float16' keeps its level data in float16)
            from . import _hip
            _hip.check(_hip.lib().fsh_ffbp_set_precision(self._form.h, 1), 'ffbp_set_precision')

    def rereference(self, coef, sign=-1):
        """Apply the phase re-reference of a CPHD read on the device as each call takes its history: sample k of
        pulse p times exp(sign 2 pi i f_k coef[p]), f_k = fmin + k df (fastsar/io.py:384-420; fastsar/cphd.py:401 moves
        the reference to the grid centre with coef = 2 dref / c, sign -1). coef None removes it."""
        from . import _hip
        if coef is not None:
            coef = np.asarray(coef, np.float64)
            if coef.shape != (self.P,) or not np.all(np.isfinite(coef)):
                raise ValueError(f'coef must hold one finite value per pulse ({self.P}), got shape {coef.shape}')
            if sign not in (-1, 1):
                raise ValueError('sign must be +1 or -1')
        if self._gate is not None:
            # a gated former re-references the whole history before its gate cuts it, on the history's own grid
            self._gate.set_ramp(None if coef is None else np.stack([sign * self._fmin * coef, sign * self._df * coef], 1))
            return
        if coef is None:
            _hip.check(_hip.lib().fsh_ffbp_set_ramp(self._form.h, None), 'ffbp_set_ramp')
            return
        ab = np.ascontiguousarray(np.stack([sign * self._fmin * coef, sign * self._df * coef], 1))
        _hip.check(_hip.lib().fsh_ffbp_set_ramp(self._form.h, ab.ctypes.data_as(ctypes.c_void_p)), 'ffbp_set_ramp')

    def adopt_buffers(self, donor):
        """Take over the device buffers of another ImageFormer whose plan has this one's shape (the formers of a
        mosaic's patches of equal shape): -> True when taken. The donor makes its own again if called later."""
        if not isinstance(donor, ImageFormer) or donor is self:
            return False
        return self._form.adopt(donor._form)

    def stage(self, S):
        """The hip former takes the history as it is: former(former.stage(S)) equals former(S). (fastsar/api.py:471-480)"""
        return S

    def __call__(self, S, out=None):
        """S [P, K] complex phase history (numpy, a fastsar_amd._hip.DeviceArray, or a 16-bit history from
        fastsar_amd.precision.pack) -> complex64 image [nx, ny]. The decode, the window and the scan for NaN and inf
        run on the device. out: an optional DeviceArray for the image."""
        from . import _hip
        from .precision import TypedHistory
        if isinstance(S, TypedHistory):
            if self._gate is not None:
                raise ValueError('a typed history is formed ungated: pass it as complex64 to a gated former')
            return self._form.form16(S, out=out)
        if not isinstance(S, _hip.DeviceArray):
            S = np.asarray(S)
            if S.shape != (self.P, self.K):
                raise ValueError(f'phase history must be {(self.P, self.K)}, got {S.shape}')
            _check_history(S, finite=False)
            S = np.ascontiguousarray(S, np.complex64)
        if self._gate is not None:
            # gated on the device (windowed and scanned for NaN and inf there) into a buffer the former keeps
            if self._gated is None:
                self._gated = _hip.DeviceArray((self.P, self.range_gate['kept']), np.complex64)
            S = self._gate(S, out=self._gated)
        return self._form(S, out=out)

    def memory(self):
        """What full speed needs on the device and what is available to this former: dict(backend, needed, available,
        full_speed, parts, group) in bytes (fastsar/api.py:643-649). needed follows the former's own buffers
        (fastsar_amd.memory.hip_full_speed: a host history, its windowed copy, the image, the group buffers at the
        default group, the plan constants; and the range gate's work buffers); available is the free device memory
        plus what the former already holds; group is the first-level group the available memory allows."""
        from . import _hip
        forms = self._form.level_forms()
        gated = self.range_gate is not None
        kw = dict(host_input=not gated, weights=self._weighted, window=bool(self.window) and not gated)
        need, parts = _mem.hip_full_speed(self._plan, forms, **kw)
        if gated:
            g = self.range_gate
            chunk = max(1, min(self.P, (1 << 30) // (8 * (g['Kp'] + g['N'] + self.K))))
            parts['range_gate'] = float(8 * chunk * (g['Kp'] + g['N'] + self.K) + 8 * self.P * g['kept'])
            need += parts['range_gate']
        free, _ = _hip.mem_info()
        avail = float(free + self._form.info()['device_bytes'])
        return dict(backend=self.backend, needed=float(need), available=avail, full_speed=need <= avail, parts=parts,
                    group=_mem.hip_group(self._plan, forms, avail - parts.get('range_gate', 0.0), **kw))


def form_image(S, ant, fmin, df, nx, ny, spx, spy, e1=(1.0, 0.0, 0.0), e2=(0.0, 1.0, 0.0), *, algorithm='ffbp',
               backend='auto', precision='float32', window=True, T='auto', levels=None, pmax=0.4, pfa_guard=300.0,
               pfa_support='union', target_db=-40.0, ref=None, interp=None, upsample=None, center=None,
               range_gate='auto'):
    """The complex image [nx, ny] (complex64). algorithm 'ffbp' (ImageFormer) or 'bp' (ExactFormer with interp,
    upsample, center); 'pfa' belongs to the reference. (fastsar/api.py:680-716)"""
    if algorithm not in ('ffbp', 'pfa', 'bp'):
        raise ValueError("algorithm must be 'ffbp', 'pfa' or 'bp'")
    if algorithm == 'pfa':
        raise ValueError("algorithm 'pfa' (polar format) belongs to the reference package; this package forms "
                         "'ffbp' and 'bp'")
    if algorithm != 'bp':
        for n, v in (('interp', interp), ('upsample', upsample), ('center', center)):
            if v is not None:
                raise ValueError(f"{n} applies to exact backprojection (algorithm='bp'), not to algorithm={algorithm!r}")
    S = _check_history(np.asarray(S), len(_check_positions(ant)), finite=False)
    if algorithm == 'bp':
        from .exact import ExactFormer
        return ExactFormer(ant, fmin, df, S.shape[1], nx, ny, spx, spy, e1, e2, backend=backend, window=window,
                           interp='cubic' if interp is None else interp, upsample=upsample, ref=ref, center=center)(S)
    return ImageFormer(ant, fmin, df, S.shape[1], nx, ny, spx, spy, e1, e2, backend=backend, precision=precision,
                       window=window, T=T, levels=levels, pmax=pmax, target_db=target_db, ref=ref,
                       range_gate=range_gate)(S)
