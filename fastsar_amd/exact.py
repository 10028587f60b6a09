"""Exact backprojection onto the form_image grid (ExactFormer, fastsar/exact.py) and onto points (backproject,
fastsar/bp.py), on the 'hip' backend.

    former = ExactFormer(ant, fmin, df, K, nx, ny, spx, spy, e1, e2)      # backend='auto'
    img = former(S)

The grid is form_image's, about the scene reference point or `center`. The host plans, once per former and in
float64: the tile size (the largest whose third-order range expansion error h^4 / (2 r^3), as a phase at f_max,
stays under 3e-4 rad), the tile centres, and the window of range bins each pulse can reach. A call range-compresses
each chunk of pulses on the device (window, centred zero padding to nfft = 2^ceil(log2(upsample K)), inverse FFT),
keeps the reachable bins and backprojects them: per tile and pulse the centre terms in float64, per pixel the
expansion in float32 and a cubic (or linear) read.

backproject(..., backend='hip') forms onto arbitrary points with exact float64 ranges and phases and complex128
profiles: the float64 path of fastsar/bp.py (its 'cpu' backend), on the device.
"""
import ctypes
import warnings

import numpy as np

from . import _hip
from .plan import C

TILES = ((32, 32), (16, 16), (8, 16), (8, 8), (4, 4))     # candidate expansion tiles, largest first
PHASE_LIMIT = 3e-4                                 # rad (fastsar/exact.py:369)


class ExactDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ('P', 'K', 'nx', 'ny', 'tx', 'ty', 'ntx', 'nty', 'nfft', 'W', 'chunk', 'cubic')] + \
               [(n, ctypes.c_double) for n in ('spx', 'spy', 'inv_dr', 'kcyc')] + \
               [('e1', ctypes.c_double * 3), ('e2', ctypes.c_double * 3),
                ('ant', _hip.P_f64), ('ref', _hip.P_f64), ('cen', _hip.P_f64), ('blo', ctypes.POINTER(ctypes.c_longlong)),
                ('wp', _hip.P_f32), ('wk', _hip.P_f32), ('profiles', ctypes.c_int)]


def _bind():
    L = _hip.lib()
    if not getattr(L, '_exact_bound', False):
        vp, ci = ctypes.c_void_p, ctypes.c_int
        L.fsh_exact_create.argtypes = [ctypes.POINTER(ExactDesc), ctypes.POINTER(vp)]
        L.fsh_exact_form.argtypes = [vp, vp, ci, vp, ci, _hip.P_i32]
        L.fsh_exact_destroy.argtypes = [vp]
        L.fsh_bp_points_f64.argtypes = [vp, ci, ci, vp, vp, vp, vp, vp, ctypes.c_long, ctypes.c_double, ctypes.c_double, ci,
                                        ci, vp, ci]
        L.fsh_sim_points.argtypes = [vp, vp, ci, ci, ctypes.c_double, ctypes.c_double, vp, vp, ci, vp, vp]
        L.fsh_add_noise.argtypes = [vp, ctypes.c_size_t, ctypes.c_double, ctypes.c_ulonglong, vp]
        L.fsh_exact_form_typed.argtypes = [vp, vp, ci, vp, ci, vp, ci, _hip.P_i32]
        L.fsh_exact_set_ramp.argtypes = [vp, vp]
        for n in ('fsh_exact_create', 'fsh_exact_form', 'fsh_exact_form_typed', 'fsh_exact_set_ramp', 'fsh_exact_destroy',
                  'fsh_bp_points_f64', 'fsh_sim_points', 'fsh_add_noise'):
            getattr(L, n).restype = ci
        L._exact_bound = True
    return L


def plane_points(nx, ny, spx, spy, e1=(1.0, 0.0, 0.0), e2=(0.0, 1.0, 0.0), height=None):
    """Points [nx, ny, 3] of the form_image grid, x_ij = (i - nx/2) spx e1 + (j - ny/2) spy e2, optionally displaced
    along the plane normal by height [nx, ny]. (fastsar/bp.py:351-361)"""
    e1, e2 = np.asarray(e1, np.float64), np.asarray(e2, np.float64)
    X, Y = np.meshgrid((np.arange(nx) - nx / 2.0) * spx, (np.arange(ny) - ny / 2.0) * spy, indexing='ij')
    p = X[..., None] * e1 + Y[..., None] * e2
    if height is not None:
        n = np.cross(e1, e2)
        n = n / np.linalg.norm(n) * np.sign(n[2] if abs(n[2]) > 1e-12 else 1.0)
        p = p + np.asarray(height, np.float64)[..., None] * n
    return p


def _integer(v, name, least):
    """v as an int when it is an integral finite number of at least `least`, else ValueError."""
    bad = isinstance(v, (bool, np.bool_)) or not np.isscalar(v)
    if not bad:
        try:
            bad = not (np.isfinite(v) and int(v) == v and v >= least)
        except (TypeError, ValueError):
            bad = True
    if bad:
        raise ValueError(f'{name} must be an integer of at least {least}, got {v!r}')
    return int(v)


def _number(v):
    try:
        return float(v)
    except (TypeError, ValueError):
        return float('nan')


def _distance_to_rectangle(ant, x0, e1, e2, a0, a1, b0, b1):
    """Per pulse the least distance from ant [P, 3] to the plane rectangle x0 + a e1 + b e2 (a in [a0, a1], b in
    [b0, b1]; e1, e2 orthonormal). The distance to a point of the plane is convex in (a, b) and smallest at the foot
    of the perpendicular, so clamping the foot's coordinates into the rectangle gives the minimum."""
    rel = ant - x0
    foot_a = np.clip(rel @ e1, a0, a1)
    foot_b = np.clip(rel @ e2, b0, b1)
    return np.linalg.norm(rel - foot_a[:, None] * e1 - foot_b[:, None] * e2, axis=1)


def _farthest_corner(ant, x0, e1, e2, a0, a1, b0, b1):
    """Per pulse the largest distance to the same rectangle: a convex function is largest at a vertex."""
    corners = np.array([x0 + a * e1 + b * e2 for a in (a0, a1) for b in (b0, b1)])
    return np.linalg.norm(corners[None] - ant[:, None], axis=2).max(1)


def expansion_tile(rmin, kmax, spx, spy, tiles=TILES, limit=PHASE_LIMIT):
    """The expansion tile: about each tile centre the range to a pixel at offset d is a series in |d| / r kept to third
    order, whose remainder is at most h^4 / (2 r^3) metres (h the tile's half-diagonal, r the least distance from the
    antenna to a tile centre, at least rmin - h), a phase of kmax = 4 pi f_max / c times that. Tiles are tried largest
    first; one counts only where r >= 4 h (the series is not used closer). -> (tx, ty, remainder in metres, phase in
    rad) of the first tile within `limit`, else of the smallest usable one; None when no tile is usable."""
    usable = []
    for tx, ty in tiles:
        h = 0.5 * np.hypot((tx - 1) * spx, (ty - 1) * spy)
        r = rmin - h
        if r >= 4 * h:
            rem = h ** 4 / (2 * r ** 3)
            usable.append((tx, ty, rem, kmax * rem))
    within = [u for u in usable if u[3] <= limit]
    return within[0] if within else (usable[-1] if usable else None)


class ExactPlan:
    """The host plan of an exact former, in float64: the expansion tile (tx, ty) and its predicted phase error, the
    tile centres cen [ntx nty, 3] (tile (i, j) at pixel (i tx + (tx - 1)/2, j ty + (ty - 1)/2) of the grid), and the
    range bins each pulse reaches: profiles of nfft = 2^ceil(log2(upsample K)) bins (zero delay at bin nfft/2, a bin
    1 / inv_dr metres), of which the W from blo[p] are kept. The bins cover the grid with one tile of margin on every
    side (the last tiles may hang past the grid), widened by the expansion's remainder and the interpolation's taps.
    Decisions as fastsar/exact.py:415-540 makes them."""

    def __init__(self, ant, ref, fmin, df, K, nx, ny, spx, spy, e1, e2, center, upsample, tiles=TILES):
        self.nfft = 1 << int(np.ceil(np.log2(upsample * K)))
        self.inv_dr = 2.0 * df * self.nfft / C
        self.kcyc = 2.0 * (fmin + (K // 2) * df) / C
        x0 = center + (-nx / 2.0) * spx * e1 + (-ny / 2.0) * spy * e2          # the centre of pixel (0, 0)
        ga, gb = (nx - 1) * spx, (ny - 1) * spy                               # the grid's extent along e1, e2
        rmin = float(_distance_to_rectangle(ant, x0, e1, e2, 0.0, ga, 0.0, gb).min())
        kmax = 4 * np.pi * (fmin + (K - 1) * df) / C
        pick = expansion_tile(rmin, kmax, spx, spy, tiles)
        if pick is None:
            raise ValueError(f'the grid comes within {rmin:.1f} m of the antenna, too close for the tiled range expansion '
                             f'of ExactFormer; use backproject on plane_points of the grid')
        self.tx, self.ty, self.expansion_m, self.phase = pick
        self.rmin = rmin
        self.predicted_error_db = float(20 * np.log10(max(self.phase, 1e-30)))
        self.ntx, self.nty = -(-nx // self.tx), -(-ny // self.ty)
        ci = (np.arange(self.ntx) * self.tx + 0.5 * (self.tx - 1))[:, None, None]
        cj = (np.arange(self.nty) * self.ty + 0.5 * (self.ty - 1))[None, :, None]
        self.cen = np.ascontiguousarray((x0 + ci * spx * e1 + cj * spy * e2).reshape(-1, 3))
        # reachable bins: the grid and a tile of margin, its nearest point and farthest corner per pulse
        ma, mb = self.tx * spx, self.ty * spy
        near = _distance_to_rectangle(ant, x0, e1, e2, -ma, ga + ma, -mb, gb + mb)
        far = _farthest_corner(ant, x0, e1, e2, -ma, ga + ma, -mb, gb + mb)
        self.slack = 4 + int(np.ceil(self.expansion_m * self.inv_dr))
        zero = self.nfft // 2
        self.blo = np.ascontiguousarray(np.floor((near - ref) * self.inv_dr).astype(np.int64) + zero - self.slack)
        last = np.ceil((far - ref) * self.inv_dr).astype(np.int64) + zero + self.slack
        self.W = int((last - self.blo).max()) + 4
        self.lo = (self.blo - zero).astype(np.int32)
        # the grid itself (no margin) must lie within one period of the profiles, c / (2 df)
        near0 = _distance_to_rectangle(ant, x0, e1, e2, 0.0, ga, 0.0, gb)
        far0 = _farthest_corner(ant, x0, e1, e2, 0.0, ga, 0.0, gb)
        self.need = int((np.ceil(far0 * self.inv_dr) - np.floor(near0 * self.inv_dr)).max()) + 2 * self.slack + 4
        self.span_m = float((far0 - near0).max())


def _point(v, what):
    """A finite 3-vector in metres, else ValueError."""
    try:
        p = np.asarray(v, np.float64)
    except (TypeError, ValueError):
        p = None
    if p is None or p.shape != (3,) or not np.isfinite(p).all():
        raise ValueError(f'{what} must be a point (a finite 3-vector) in meters, got {v!r}')
    return p


def _ranges(v, P):
    """One finite range per pulse in metres, else ValueError."""
    try:
        r = np.asarray(v, np.float64)
    except (TypeError, ValueError):
        raise ValueError('ref must be a range per pulse in meters') from None
    if r.shape != (P,):
        raise ValueError(f'ref must hold one range per pulse, shape ({P},), got {r.shape}')
    if not np.isfinite(r).all():
        raise ValueError(f'ref has non-finite values (NaN or inf), the first in pulse {int(np.argmax(~np.isfinite(r)))}')
    return r


class ExactFormer:
    """Exact backprojection onto the grid of form_image, set up once for a collection geometry. former(S) ->
    complex64 image [nx, ny]. interp 'cubic' (default) or 'linear'; upsample the range oversampling (default 8);
    chunk the pulses range compressed at a time; ref [P] the reference ranges (default |ant|); center the grid's
    centre (default the origin). former.tile (the expansion tile) and former.predicted_error_db as the reference's
    (fastsar/exact.py:404-540). A former must not be called from two threads at once."""

    def __init__(self, ant, fmin, df, K, nx, ny, spx, spy, e1=(1.0, 0.0, 0.0), e2=(0.0, 1.0, 0.0), *, backend='auto',
                 window=True, interp='cubic', upsample=None, ref=None, center=None, chunk=1024, profiles='auto'):
        from .api import _backend, _check_positions, _check_grid, _window
        if interp not in ('cubic', 'linear'):
            raise ValueError(f"interp must be 'cubic' or 'linear', got {interp!r}")
        if profiles not in ('auto', 'float32', 'float16', 'float16-pair'):
            raise ValueError(f"profiles must be 'auto', 'float32', 'float16' or 'float16-pair', got {profiles!r}")
        if profiles == 'float16-pair' and interp != 'linear':
            raise ValueError("profiles='float16-pair' applies to interp='linear'")
        # 'auto' keeps float32 profiles on every grid: the region kernel reads its windows from LDS, so float16
        # storage saves only the staging reads and adds the conversions
        self.profiles = 'float32' if profiles == 'auto' else profiles
        self.ant = _check_positions(ant)
        self.P = len(self.ant)
        self.K = _integer(K, 'K (frequency samples per pulse)', 2)
        fmin_v, df_v = _number(fmin), _number(df)
        if not (np.isfinite(fmin_v) and np.isfinite(df_v) and df_v > 0):
            raise ValueError(f'fmin must be a finite frequency and df a positive frequency spacing in Hz, got {fmin} and {df}')
        nx, ny, e1, e2 = _check_grid(nx, ny, spx, spy, e1, e2)
        self.nx, self.ny, self.spx, self.spy = int(nx), int(ny), float(spx), float(spy)
        self.e1, self.e2 = e1, e2
        self.center = np.zeros(3) if center is None else _point(center, 'center')
        self.ref = np.linalg.norm(self.ant, axis=1) if ref is None else _ranges(ref, self.P)
        self.cubic = interp == 'cubic'
        self.upsample = 8 if upsample is None else _integer(upsample, 'upsample', 1)
        self.chunk = _integer(chunk, 'chunk', 1)
        self.backend = _backend(backend)
        self.fmin, self.df, self.window = fmin_v, df_v, window
        self.wp, self.wk = _window(self.P, self.K) if window else (None, None)
        plan = ExactPlan(self.ant, self.ref, fmin_v, df_v, self.K, self.nx, self.ny, self.spx, self.spy, self.e1, self.e2,
                         self.center, self.upsample)
        self.plan = plan
        self.nfft, self.inv_dr, self.kcyc = plan.nfft, plan.inv_dr, plan.kcyc
        self.tx, self.ty, self.tile = plan.tx, plan.ty, (plan.tx, plan.ty)
        self._expansion_m, self.predicted_error_db = plan.expansion_m, plan.predicted_error_db
        self.ntx, self.nty, self.cen = plan.ntx, plan.nty, plan.cen
        self.blo, self.lo, self.W = plan.blo, plan.lo, plan.W
        if plan.phase > PHASE_LIMIT:
            warnings.warn(f'ExactFormer: at {plan.rmin:.0f} m from the antenna, even {self.tx} by {self.ty} pixel tiles leave '
                          f'a predicted phase error of {plan.phase:.1e} rad ({self.predicted_error_db:.1f} dB) from the range '
                          'expansion; backproject on plane_points of the grid is exact at any range',
                          UserWarning, stacklevel=2)
        if plan.need > self.nfft:
            raise ValueError(f'the grid spans up to {plan.span_m:.1f} m of range from one pulse, more than '
                             f'this collection resolves without ambiguity (c / (2 df) = {C / (2 * df_v):.1f} m less the '
                             f'interpolation margin: the grid needs {plan.need} of the {self.nfft} range bins at '
                             f'upsample={self.upsample}); form a smaller grid (or several) or use samples spaced more '
                             'finely in frequency')
        self._h = None

    def _handle(self):
        if self._h is None:
            L = _bind()
            d = ExactDesc()
            d.P, d.K, d.nx, d.ny, d.tx, d.ty = self.P, self.K, self.nx, self.ny, self.tx, self.ty
            d.ntx, d.nty, d.nfft, d.W, d.chunk, d.cubic = self.ntx, self.nty, self.nfft, self.W, min(self.chunk, self.P), int(self.cubic)
            d.spx, d.spy, d.inv_dr, d.kcyc = self.spx, self.spy, self.inv_dr, self.kcyc
            d.profiles = {'float32': 0, 'float16': 1, 'float16-pair': 2}[self.profiles]
            for k in range(3):
                d.e1[k], d.e2[k] = self.e1[k], self.e2[k]
            keep = [np.ascontiguousarray(self.ant), np.ascontiguousarray(self.ref), self.cen, self.blo]
            d.ant = keep[0].ctypes.data_as(_hip.P_f64)
            d.ref = keep[1].ctypes.data_as(_hip.P_f64)
            d.cen = keep[2].ctypes.data_as(_hip.P_f64)
            d.blo = keep[3].ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
            if self.wp is not None:
                keep += [np.ascontiguousarray(self.wp, np.float32), np.ascontiguousarray(self.wk, np.float32)]
                d.wp, d.wk = keep[-2].ctypes.data_as(_hip.P_f32), keep[-1].ctypes.data_as(_hip.P_f32)
            h = ctypes.c_void_p()
            _hip.check(L.fsh_exact_create(ctypes.byref(d), ctypes.byref(h)), 'exact_create')
            self._h = h
        return self._h

    def rereference(self, coef, sign=-1):
        """Apply the phase re-reference of a CPHD read on the device as each call takes its history: sample k of
        pulse p times exp(sign 2 pi i f_k coef[p]), f_k = fmin + k df (fastsar/io.py:384-420; fastsar/cphd.py:401 moves
        the reference to the grid centre with coef = 2 dref / c, sign -1). coef None removes it."""
        h = self._handle()
        if coef is None:
            _hip.check(_bind().fsh_exact_set_ramp(h, None), 'exact_set_ramp')
            return
        coef = np.asarray(coef, np.float64)
        if coef.shape != (self.P,) or not np.all(np.isfinite(coef)):
            raise ValueError(f'coef must hold one finite value per pulse ({self.P}), got shape {coef.shape}')
        if sign not in (-1, 1):
            raise ValueError('sign must be +1 or -1')
        ab = np.ascontiguousarray(np.stack([sign * self.fmin * coef, sign * self.df * coef], 1))
        _hip.check(_bind().fsh_exact_set_ramp(h, ab.ctypes.data_as(ctypes.c_void_p)), 'exact_set_ramp')

    def memory(self):
        """dict(backend, needed, available, full_speed, parts) in bytes. (fastsar/exact.py:736-761)"""
        ch = min(self.chunk, self.P)
        parts = dict(history=self.P * self.K * 8, padded=ch * self.nfft * 8, fft=ch * self.nfft * 8,
                     window=ch * self.W * {'float32': 8, 'float16': 12, 'float16-pair': 16}[self.profiles],
                     image=self.nx * self.ny * 8, geometry=self.cen.nbytes + self.P * 40)
        need = float(sum(parts.values()))
        avail = float(_hip.device_info()['mem'])
        return dict(backend=self.backend, needed=need, available=avail, full_speed=need <= avail, parts=parts)

    def __call__(self, S, out=None):
        """S [P, K] complex phase history (numpy, a DeviceArray, or a 16-bit history from
        fastsar_amd.precision.pack) -> complex64 image [nx, ny] (or into out, a DeviceArray). NaN and inf are found
        on the device and raise ValueError naming the first pulse."""
        from .api import _check_history
        from .precision import TypedHistory
        if isinstance(S, TypedHistory):
            if S.shape != (self.P, self.K):
                raise ValueError(f'phase history must be {(self.P, self.K)}, got {S.shape}')
            img, img_dev = _hip.output(out, (self.nx, self.ny))
            bad = ctypes.c_int(-1)
            h = self._handle()
            data = S.data if S.on_device else np.ascontiguousarray(S.data)
            inv = S.inv if (S.on_device or S.inv is None) else np.ascontiguousarray(S.inv, np.float32)
            _hip.check(_bind().fsh_exact_form_typed(h, _hip.ptr(data), S.code, _hip.ptr(inv), int(S.on_device),
                                                    _hip.ptr(img), int(img_dev), ctypes.byref(bad)), 'exact_form')
            return img
        on_dev = isinstance(S, _hip.DeviceArray)
        if not on_dev:
            S = _check_history(np.asarray(S), self.P, finite=False)
            if S.shape[1] != self.K:
                raise ValueError(f'phase history must have {self.K} samples per pulse, got {S.shape[1]}')
            S = np.ascontiguousarray(S, np.complex64)
        elif S.shape != (self.P, self.K) or S.dtype != np.complex64:
            raise ValueError(f'phase history must be complex64 {(self.P, self.K)}')
        img, img_dev = _hip.output(out, (self.nx, self.ny))
        h = self._handle()
        bad = ctypes.c_int(-1)
        _hip.check(_bind().fsh_exact_form(h, _hip.ptr(S), int(on_dev), _hip.ptr(img), int(img_dev), ctypes.byref(bad)),
                   'exact_form')
        return img

    def __del__(self):
        try:
            if self._h is not None and _hip._lib is not None:
                _hip._lib.fsh_exact_destroy(self._h)
        except Exception:
            pass


def backproject(S, ant, fmin, df, points, *, ref=None, backend='auto', upsample=8, window=True, chunk=256,
                interp='linear'):
    """Complex image at points [..., 3] (complex128, shape [...]) by exact float64 backprojection: profiles
    oversampled `upsample` times (nfft a power of two) in complex128, read linearly (or cubic) at the exact float64
    range of every point and pulse. ref [P]: the reference ranges (default |ant|). Monostatic.
    (fastsar/bp.py:303-348, its float64 'cpu' path)"""
    from .api import _backend, _check_history, _check_positions, _window
    _backend(backend)
    S = _check_history(np.asarray(S))
    P, K = S.shape
    tx = _check_positions(ant, 'antenna positions', P)
    points = np.asarray(points, np.float64)
    if points.ndim < 1 or points.shape[-1] != 3 or points.size == 0:
        raise ValueError(f'points must be [..., 3] with at least one point, got shape {points.shape}')
    ref = np.linalg.norm(tx, axis=1) if ref is None else np.asarray(ref, np.float64)
    if ref.shape != (P,):
        raise ValueError(f'ref must hold one range per pulse, shape ({P},), got {ref.shape}')
    shape = points.shape[:-1]
    nfft = 1 << int(np.ceil(np.log2(upsample * K)))
    wp, wk = _window(P, K) if window else (None, None)
    dS = _hip.DeviceArray.from_host(np.ascontiguousarray(S, np.complex64))
    return bp_points_f64_device(dS, tx, ref, fmin, df, points.reshape(-1, 3), nfft, interp == 'cubic', wp, wk, chunk).reshape(shape)


def bp_points_f64_device(dS, ant, ref, fmin, df, pts, nfft, cubic=False, wp=None, wk=None, chunk=256):
    """fsh_bp_points_f64 on a device history dS [P, K] complex64 -> complex128 [N] on the host."""
    L = _bind()
    P, K = dS.shape
    dant = _hip.DeviceArray.from_host(np.ascontiguousarray(ant, np.float64))
    dref = _hip.DeviceArray.from_host(np.ascontiguousarray(ref, np.float64))
    dpts = _hip.DeviceArray.from_host(np.ascontiguousarray(pts, np.float64))
    out = _hip.DeviceArray((len(pts),), np.complex128).zero()
    dwp = None if wp is None else _hip.DeviceArray.from_host(np.ascontiguousarray(wp, np.float32))
    dwk = None if wk is None else _hip.DeviceArray.from_host(np.ascontiguousarray(wk, np.float32))
    _hip.check(L.fsh_bp_points_f64(_hip.ptr(dS), P, K, _hip.ptr(dwp), _hip.ptr(dwk), _hip.ptr(dant), _hip.ptr(dref),
                                   _hip.ptr(dpts), len(pts), float(fmin), float(df), int(nfft), int(bool(cubic)), _hip.ptr(out),
                                   int(chunk)), 'bp_points_f64')
    return out.to_host()


def simulate_device(col, pos, amp, ref=None, clutter_sigma=None, seed=0):
    """The phase history of a point scene on the device (fsh_sim_points: float64 ranges and phases, float64 sums,
    stored complex64) plus optional white clutter: a DeviceArray [P, K] complex64."""
    L = _bind()
    P, K = col.ant.shape[0], col.K
    ant = np.ascontiguousarray(col.ant, np.float64)
    ref = np.linalg.norm(ant, axis=1) if ref is None else np.asarray(ref, np.float64)
    S = _hip.DeviceArray((P, K), np.complex64).zero()
    dant, dref = _hip.DeviceArray.from_host(ant), _hip.DeviceArray.from_host(np.ascontiguousarray(ref))
    dpos = _hip.DeviceArray.from_host(np.ascontiguousarray(pos, np.float64))
    damp = _hip.DeviceArray.from_host(np.ascontiguousarray(np.asarray(amp, np.complex128)))
    _hip.check(L.fsh_sim_points(_hip.ptr(dant), _hip.ptr(dref), P, K, float(col.fmin), float(col.df), _hip.ptr(dpos),
                                _hip.ptr(damp), len(pos), _hip.ptr(S), None), 'sim_points')
    if clutter_sigma:
        _hip.check(L.fsh_add_noise(_hip.ptr(S), P * K, float(clutter_sigma), int(seed), None), 'add_noise')
    _hip.check(L.fsh_synchronize(), 'synchronize')
    return S
