"""Patch mosaics: a planar grid formed patch by patch, each patch by an ImageFormer from the pulses that see it, its
history re-referenced to the patch centre and cut in range to the patch on the device. (fastsar/patches.py:619-970)

    img = form_mosaic(fx, ant, origin, nx, ny, spx, spy, e1, e2, patch=(128, 128), beam=beam, awin=awin)

Phase history: the frequency-domain form with per-pulse reference ranges, fx = dict(S [P, K], fmin, df, ref [P]), in
which a scatterer at x adds exp(-j 4 pi f_k (|x - a_p| - ref_p) / c) at f_k = fmin + k df (the reference's FX
convention, fastsar/patches.py:9-16). Grid: pixel (i, j) at origin + i spx e1 + j spy e2 (origin is pixel (0, 0)).

Each tile of patch = (mx, my) pixels is formed as a patch of (mx + 2 crop, my + 2 crop) pixels centred on it, and
crop pixels are dropped on each side. Per patch, on the host: its centre c and a 9 x 9 lattice over it; the pulses
that serve it, one contiguous run [lo, hi) (all pulses, `pulses`, or those whose beam coordinate |u| <= umax at some
lattice point); positions of dropped pulses filled in on the line between their neighbours (fill_gaps), so that the
pulse axis keeps an even spacing for the levels' pulse filters. On the device, in the former's one call: every pulse
moved from ref_p to |a_p - c| (a phase ramp linear in frequency), then, where the patch spans less than the swath,
the range gate (fastsar_amd.gate) keeps only the range bins the patch reaches, and the levels and the final stage
form the patch. With an azimuth window awin the pulses are weighted per pixel by awin(u / umax), applied in the
final stage (ImageFormer's aperture_weight).

Patches of equal shape share their plan's shape: formers are kept per plan signature (plan.plan_signature) up to
`keep` at a time, and the former of a patch whose signature, pulse count and gate match an earlier one takes over
that former's group buffers, history buffer and pinned staging ring (ImageFormer.adopt_buffers, fsh_ffbp_adopt), so
that only its plan constants are made anew; the reference keeps its compiled programs per signature the same way
(fastsar/api.py:661-677).
"""
import numpy as np

from .plan import C


def fill_gaps(a):
    """Positions a [n, 3] with the pulses a platform dropped put back: where consecutive positions are m >= 2 median
    steps apart, m - 1 positions evenly between them. -> (positions [n', 3], index [n] of each given pulse among
    them). Zero pulses at the inserted positions add nothing to the image and keep the pulse spacing even, which the
    levels' pulse filters assume. (fastsar/patches.py:571-590)"""
    a = np.asarray(a, np.float64)
    if len(a) < 2:
        return a, np.arange(len(a))
    gaps = np.linalg.norm(a[1:] - a[:-1], axis=1)
    unit = float(np.median(gaps))
    if not unit > 0:
        return a, np.arange(len(a))
    count = np.maximum(np.rint(gaps / unit).astype(np.int64), 1)          # steps each gap spans
    if count.max() < 2:
        return a, np.arange(len(a))
    where = np.concatenate([[0], np.cumsum(count)])                        # the given pulses' new indices
    out = np.empty((int(where[-1]) + 1, 3))
    out[where] = a
    for g in np.flatnonzero(count > 1):
        frac = (np.arange(1, count[g]) / count[g])[:, None]
        out[where[g] + 1:where[g + 1]] = a[g] + frac * (a[g + 1] - a[g])
    return out, where


def beam_span(beam, P, pts, umax, step=64):
    """The pulses [lo, hi) that see any of the points pts [m, 3], |beam(idx, pts)| <= umax, or None. A point's pulses
    are contiguous, so the beam is evaluated on every step-th pulse and then pulse by pulse about the first and the
    last pulse found. (fastsar/patches.py:551-568)"""
    def sees(idx):
        return (np.abs(beam(np.asarray(idx), pts)) <= umax).any(axis=1)
    grid = np.unique(np.concatenate([np.arange(0, P, step), [P - 1]]))
    hits = np.flatnonzero(sees(grid))
    if hits.size == 0:
        every = np.flatnonzero(sees(np.arange(P)))
        return None if every.size == 0 else (int(every[0]), int(every[-1]) + 1)
    left = np.arange(int(grid[hits[0] - 1]) if hits[0] > 0 else 0, int(grid[hits[0]]) + 1)
    right = np.arange(int(grid[hits[-1]]), (int(grid[hits[-1] + 1]) if hits[-1] + 1 < grid.size else P - 1) + 1)
    lo = int(left[np.flatnonzero(sees(left))[0]])
    hi = int(right[np.flatnonzero(sees(right))[-1]]) + 1
    return lo, hi


def patch_levels(n, T, most=3, longest=None):
    """Levels for a patch of n pixels (its shorter side) in tiles of T: as many as split ceil(n / T) tiles by at least
    2 per level without padding, at most `most`, and enough for the longer side (`longest` pixels) to split with
    factors of at most 16 per level. (fastsar/patches.py:162-171)"""
    tiles = -(-n // T)
    nlev = int(np.clip(np.floor(np.log2(max(tiles, 1))), 1, most))
    if longest is not None:
        while nlev < most and 16 ** nlev < -(-longest // T):
            nlev += 1
    return nlev


class _FormerCache:
    """Formers of a mosaic kept per plan signature (at most `keep`, the oldest dropped first)."""

    def __init__(self, keep):
        self.keep, self.items = keep, {}

    def take(self, key):
        return self.items.pop(key, None)

    def give(self, key, former):
        self.items[key] = former
        while len(self.items) > self.keep:
            self.items.pop(next(iter(self.items)))


def form_mosaic(fx, ant, origin, nx, ny, spx, spy, e1=(1.0, 0.0, 0.0), e2=(0.0, 1.0, 0.0), patch=(128, 128), crop=0,
                beam=None, umax=1.0, awin=None, pulses=None, backend='auto', T='auto', levels='auto', target_db=-40.0,
                info=None, precision='float32', range_gate=True, keep=4):
    """Complex image [nx, ny] (complex64) on the grid origin + i spx e1 + j spy e2 from the FX phase history fx and the
    antenna positions ant [P, 3] of any track, formed patch by patch with ImageFormer on the hip backend.

    beam(idx, points [m, 3]) -> u [len(idx), m]: the normalized azimuth coordinate of pulses idx at the points; a pulse
    serves a patch where |u| <= umax at one of its 9 x 9 lattice points. awin(u) (with beam): the azimuth window,
    pulse p weighted at pixel x by awin(u_p(x) / umax) with u clipped to [-umax, umax]. pulses: (lo, hi) or a
    function of the patch centre returning (lo, hi), used without a beam (default: every pulse). T, levels,
    target_db, precision: as ImageFormer's ('auto' levels: patch_levels). range_gate: cut each patch's history in
    range before formation (where it pays, ImageFormer's rule). keep: formers kept for reuse by plan signature
    (0: none; every patch makes its own buffers).
    info: a list that gets one dict per patch (tile, centre, pulses, T, levels, predicted error, reused).
    (fastsar/patches.py:619-970)"""
    from . import _hip
    from .api import ImageFormer, _backend, _check_grid, _check_positions
    from .plan import plan_signature
    _backend(backend)
    S_all = fx['S']
    fmin, df, ref = float(fx['fmin']), float(fx['df']), np.asarray(fx['ref'], np.float64)
    ant = _check_positions(ant, P=S_all.shape[0])
    if ref.shape != (len(ant),):
        raise ValueError(f"fx['ref'] must hold one range per pulse ({len(ant)}), got shape {ref.shape}")
    if beam is None and awin is not None:
        raise ValueError('awin weights pulses by their beam coordinate: give beam with it')
    nx, ny, e1, e2 = _check_grid(nx, ny, spx, spy, e1, e2)
    origin = np.asarray(origin, np.float64)
    mx, my = int(patch[0]), int(patch[1])
    crop = int(crop)
    if mx < 1 or my < 1 or crop < 0:
        raise ValueError(f'patch must be two positive pixel counts and crop >= 0, got {patch!r}, {crop!r}')
    px, py = mx + 2 * crop, my + 2 * crop
    P, K = S_all.shape
    out = np.zeros((nx, ny), np.complex64)
    cache = _FormerCache(int(keep)) if int(keep) > 0 else None
    lat = np.linspace(0.0, 1.0, 9)
    for i0 in range(0, nx, mx):
        for j0 in range(0, ny, my):
            # the patch: pixels (i0 - crop + a, j0 - crop + b), a < px, b < py; its centre in the former's convention
            # (pixel (a, b) at (a - px/2) spx e1 + (b - py/2) spy e2 about it)
            c = origin + (i0 - crop + px / 2.0) * spx * e1 + (j0 - crop + py / 2.0) * spy * e2
            corner = origin + (i0 - crop) * spx * e1 + (j0 - crop) * spy * e2
            pts = (corner + (lat[:, None, None] * (px - 1) * spx) * e1 + (lat[None, :, None] * (py - 1) * spy) * e2).reshape(-1, 3)
            if beam is not None:
                span = beam_span(beam, P, pts, umax)
            elif pulses is None:
                span = (0, P)
            else:
                span = tuple(pulses(c)) if callable(pulses) else tuple(pulses)
            if span is None or span[1] - span[0] < 2:
                continue
            lo, hi = int(span[0]), int(span[1])
            track, at = fill_gaps(ant[lo:hi] - c)
            n = len(track)
            # the re-reference of each recorded pulse from ref_p to its range to the centre, as cycles per frequency
            rc = np.linalg.norm(ant[lo:hi] - c, axis=1)
            coef = np.zeros(n)
            coef[at] = 2.0 * (ref[lo:hi] - rc) / C
            Sp = np.zeros((n, K), np.complex64)
            Sp[at] = S_all[lo:hi]
            weight = None
            if awin is not None:
                def weight(q, pul, c=c, lo=lo, at=at):
                    # pulses of the filled track -> the nearest recorded pulse; points relative to the patch centre
                    k = np.clip(np.searchsorted(at, pul), 0, len(at) - 1)
                    u = np.clip(beam(lo + k, np.asarray(q) + c) / umax, -1.0, 1.0)
                    return awin(u)
            nlev = None
            Tp = T
            if T == 'auto':
                from .plan import choose_T
                Tp, _ = choose_T(track, fmin + K * df, px, py, spx, spy, e1, e2, target_db, choices=(32, 16, 8, 4))
            if levels == 'auto':
                nlev = patch_levels(min(px, py), Tp, longest=max(px, py))
            else:
                nlev = int(levels)
            former = ImageFormer(track, fmin, df, K, px, py, spx, spy, e1, e2, backend='hip', precision=precision,
                                 window=False, T=Tp, levels=nlev, target_db=target_db, aperture_weight=weight,
                                 range_gate=bool(range_gate))
            key = (plan_signature(former._plan), n, K, former.range_gate is not None, weight is not None, precision)
            donor = cache.take(key) if cache is not None else None
            reused = donor is not None and former.adopt_buffers(donor)
            former.rereference(coef)
            img = former(Sp)
            if cache is not None:
                cache.give(key, former)
            a1, b1 = min(mx, nx - i0), min(my, ny - j0)
            out[i0:i0 + a1, j0:j0 + b1] = img[crop:crop + a1, crop:crop + b1]
            if info is not None:
                info.append(dict(tile=(i0, j0), center=c.tolist(), pulses=(lo, hi), filled=n - (hi - lo), T=former.T,
                                 levels=former.levels, predicted_error_db=former.predicted_error_db,
                                 gated=former.range_gate is not None, reused=bool(reused)))
    _hip.check(_hip.lib().fsh_synchronize())
    return out
