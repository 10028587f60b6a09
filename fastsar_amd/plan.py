"""Host-side planner of factorized backprojection, in float64 NumPy.

Factorized backprojection (FFBP) forms an image by recursive division of the output grid into tiles. At every
level each parent tile's data, held as [pulses x frequency samples], is re-referenced (rotated in phase) to the
centre of each of its children and low-pass filtered and decimated along both axes with a Kaiser-windowed sinc.
After the last level each T x T tile holds a short [Pf x Qf] block that the final stage turns into pixels with a
separable plane-wave model plus a per-pixel quadratic phase.

This module decides everything of that recursion that does not depend on the phase history:

  choose_splits      the per-level split factors of each image axis           (fastsar/ffbp2.py:34-62)
  decimator_fir      the decimation filter of one axis of one level            (fastsar/ffbp2.py:103-137)
  positions          the antenna track on a level's fractional pulse grid     (fastsar/ffbp.py:122-125)
  make_plan          tiles, decimations, filters, child centres per level     (fastsar/ffbp2.py:192-250)
  final_phase_error  the predicted error of the final stage's tile model      (fastsar/ffbp2.py:276-297)
  choose_T           the largest final tile meeting an error target           (fastsar/ffbp2.py:300-307)
  collection_arrays  the per-collection rotation phases of the first levels   (fastsar/ffbp2.py:313-342)
  pulse_plan, plan_signature                                                  (fastsar/ffbp2.py:253-273)

The plan is a plain dict with the same keys and meaning as the reference's, so that the images of every backend
agree. The arithmetic of the filter design keeps the reference's operation order, which makes the float64 taps
equal to the bit (and so the float32 taps every kernel reads).
"""
import functools
import math

import numpy as np

C = 299792458.0          # speed of light, m/s (fastsar/ffbp.py:30)
FACTORS = (8, 7, 6, 5, 4, 3, 2)
HOST_LEVELS = 2          # levels whose rotation phases are computed on the host in float64 (fastsar/ffbp2.py:310)


def _prod(v):
    out = 1
    for x in v:
        out *= int(x)
    return out


def choose_splits(n, T, nlev):
    """Split factors of one image axis, one per level, for n pixels in final tiles of T: the smallest product that
    is at least ceil(n / T), written as nlev factors from FACTORS in non-increasing order; among equal products the
    lexicographically largest tuple. When no such product exists, the same number of factors between 9 and 16.
    (fastsar/ffbp2.py:34-62)"""
    need = -(-int(n) // int(T))
    best = None
    # every non-increasing tuple of nlev factors, enumerated by depth-first search
    stack = [((), 1)]
    while stack:
        prefix, prod = stack.pop()
        if len(prefix) == nlev:
            if prod >= need:
                key = (prod, prefix)
                if best is None or prod < best[0] or (prod == best[0] and prefix > best[1]):
                    best = key
            continue
        for f in FACTORS:
            if prefix and f > prefix[-1]:
                continue
            stack.append((prefix + (f,), prod * f))
    if best is not None:
        return best[1]
    # too few levels for factors up to 8: equal factors f (the smallest f <= 16 with f^nlev >= need, taken from 16
    # downward so that the last candidate tried wins), the last factor lowered while the product still suffices
    found = None
    for f in range(16, 8, -1):
        if f ** nlev >= need:
            v = [f] * nlev
            while v and _prod(v[:-1]) * (v[-1] - 1) >= need and v[-1] > 2:
                v[-1] -= 1
            found = tuple(v)
    if found is None:
        raise ValueError(f'{nlev} levels cannot split {n} pixels into tiles of {T}: use more levels or a larger T')
    return found


@functools.lru_cache(maxsize=512)
def _design(n_in, D, pass_frac, atten):
    """The decimation filter as one kernel (see decimator_fir); memoized, read-only taps."""
    dw = (2.0 - 2.0 * pass_frac) * np.pi / D                      # transition width, rad per input sample
    half = math.ceil((atten - 8.0) / (2.285 * dw)) / 2.0 + 1.0   # Kaiser's length formula, as a half width
    m = int(math.ceil(half / D))                                  # outputs past each end of the input
    n_out = -(-n_in // D) + 2 * m
    beta = 0.1102 * (atten - 8.7)                                 # Kaiser's beta for atten > 50 dB

    def weight(d):
        # windowed sinc at distance d (input samples) from an output's centre; zero beyond half
        return np.where(np.abs(d) <= half,
                        np.sinc(d / D) * (np.i0(beta * np.sqrt(np.clip(1.0 - (d / half) ** 2, 0.0, 1.0))) / np.i0(beta)),
                        0.0)

    jc = n_out // 2                                               # a central output, far from either edge
    cj = D * (jc - m) + (D - 1) / 2.0                             # its centre on the input axis
    i = np.arange(max(0, int(math.floor(cj - half))), min(n_in, int(math.ceil(cj + half)) + 1))
    v = weight(i - cj)
    keep = v != 0
    i, v = i[keep], v[keep]
    # each input's total weight over all outputs (the decimation matrix is normalized by its row sums)
    centres = D * (np.arange(n_out) - m) + (D - 1) / 2.0
    reach = np.abs(centres[None, :] - i[:, None]) <= half
    rows = np.where(reach, weight(i[:, None] - centres[None, :]), 0.0).sum(1)
    kern = v / rows
    lo = int(i[0]) - D * (jc - m)
    L = int(i[-1] - i[0] + 1)
    pl = D * m - lo
    if pl < 0 or L != len(kern):
        raise AssertionError('decimation kernel with interior zeros or negative left padding')
    kern.flags.writeable = False
    pr = max(D * (n_out - 1) + L - pl - n_in, 0)
    return (('kern', kern), ('pl', int(pl)), ('pr', int(pr)), ('n_out', int(n_out)), ('L', int(L))), m


def decimator_fir(n_in, D, pass_frac, atten=70.0):
    """Kaiser-windowed sinc decimation of n_in samples by D as one shift-invariant kernel:
    out[j] = sum_{r < L} kern[r] * x[D j + r - pl], x zero outside [0, n_in), for j < n_out = ceil(n_in / D) + 2 m.
    Output j is centred on input D (j - m) + (D - 1) / 2; pass_frac is the passband edge as a fraction of the
    output Nyquist rate, the stopband starts at 2 - pass_frac; every input's weights over the outputs sum to one.
    Returns (dict(kern, pl, pr, n_out, L), m). (fastsar/ffbp2.py:103-137, fastsar/ffbp.py:57-78)"""
    d, m = _design(int(n_in), int(D), float(pass_frac), float(atten))
    return dict(d), m


def decimator_matrix(n_in, D, pass_frac, atten=70.0):
    """The dense decimation matrix F [n_in, n_out] (column j: the kernel placed at output j), for tests and the
    small dense forms; identity when D == 1. Each row sums to one. (fastsar/ffbp.py:57-78)"""
    if D == 1:
        return np.eye(n_in), 0
    fr, m = decimator_fir(n_in, D, pass_frac, atten)
    F = np.zeros((n_in, fr['n_out']))
    for j in range(fr['n_out']):
        for r in range(fr['L']):
            k = D * j + r - fr['pl']
            if 0 <= k < n_in:
                F[k, j] = fr['kern'][r]
    return F, m


def _margin(n_in, D, pass_frac, atten):
    """(m, n_out) of a decimation, without designing the filter (D == 1: no margin)."""
    if D == 1:
        return 0, int(n_in)
    dw = (2.0 - 2.0 * pass_frac) * np.pi / D
    half = math.ceil((atten - 8.0) / (2.285 * dw)) / 2.0 + 1.0
    m = int(math.ceil(half / D))
    return m, -(-int(n_in) // D) + 2 * m


def positions(ant, idx):
    """Antenna positions at fractional pulse indices idx: linear interpolation between neighbouring pulses, and
    linear extrapolation beyond either end. (fastsar/ffbp.py:122-125)"""
    ant = np.asarray(ant, np.float64)
    i0 = np.clip(np.floor(idx).astype(np.int64), 0, len(ant) - 2)
    return ant[i0] + (np.asarray(idx, np.float64) - i0)[:, None] * (ant[i0 + 1] - ant[i0])


def make_plan(col, nx, ny, spx, spy, T=32, nlev=3, pmax=0.4, atten=70.0, splits=None, e1=(1.0, 0.0, 0.0),
              e2=(0.0, 1.0, 0.0), pulse=None):
    """Tile grid, decimation factors and filters of every level for an nx x ny image of spacings spx, spy along the
    orthonormal axes e1, e2, centred on the origin of the collection's coordinates. col: anything with fmin, df,
    K, ant [P, 3] (sim.Collect). pulse: [(Dp, pass_p)] per level in place of the ones this track gives.
    (fastsar/ffbp2.py:192-250)

    Per level, for child tiles of hx x hy half extents and the track's unit vectors u_p:
      rk  = hx max|u.e1| + hy max|u.e2|                  half extent of a child in range (m)
      dop = hx max|diff u.e1| + hy max|diff u.e2|        its range change per pulse step (m)
      Dk  = floor(pmax (c / 2df) / 2 / rk),  Dp = floor(pmax / (2 (2 fmax / c) dop))
    with pass bands rounded up to 1/64 of the output Nyquist rate, and no decimation of an axis shorter than its
    kernel. Each level's antenna track is the previous one at the output pulses' fractional indices."""
    e1, e2 = np.asarray(e1, np.float64), np.asarray(e2, np.float64)
    sxs = choose_splits(nx, T, nlev) if splits is None else tuple(int(s[0]) for s in splits)
    sys_ = choose_splits(ny, T, nlev) if splits is None else tuple(int(s[1]) for s in splits)
    Nx, Ny = T * _prod(sxs), T * _prod(sys_)
    ox, oy = (Nx - nx) // 2, (Ny - ny) // 2
    ant = np.asarray(col.ant, np.float64)
    f0, df, K, P = float(col.fmin), float(col.df), int(col.K), int(ant.shape[0])
    ref = np.zeros((1, 3))
    i0 = np.zeros((1, 2))
    cx, cy = Nx, Ny
    out = dict(levels=[], T=T, nx=nx, ny=ny, Nx=Nx, Ny=Ny, ox=ox, oy=oy, spx=spx, spy=spy,
               split=list(zip(sxs, sys_)), e1=e1, e2=e2)
    for n, (sx, sy) in enumerate(zip(sxs, sys_)):
        cx, cy = cx // sx, cy // sy
        gi, gj = np.meshgrid(np.arange(sx) * cx, np.arange(sy) * cy, indexing='ij')
        ci0 = i0[:, None, :] + np.stack([gi.ravel(), gj.ravel()], axis=1)[None, :, :]          # [B, C, 2] first pixel
        cen = (((ci0[..., 0] + (cx - 1) / 2.0 - ox - nx / 2.0) * spx)[..., None] * e1 +
               ((ci0[..., 1] + (cy - 1) / 2.0 - oy - ny / 2.0) * spy)[..., None] * e2)          # [B, C, 3] centres
        u = ant / np.linalg.norm(ant, axis=1)[:, None]
        u1, u2 = u @ e1, u @ e2
        hx, hy = 0.5 * cx * spx, 0.5 * cy * spy
        rk = hx * np.abs(u1).max() + hy * np.abs(u2).max()
        dop = hx * np.abs(np.diff(u1)).max() + hy * np.abs(np.diff(u2)).max()
        kfar = 2.0 * (f0 + K * df) / C                                                          # cycles per m at fmax
        Dk = max(1, int(pmax * (C / (2.0 * df) / 2.0) / rk))
        Dp = max(1, int(pmax / (2.0 * kfar * dop)))
        pass_k = math.ceil(64 * rk / (C / (2.0 * df * Dk) / 2.0)) / 64
        pass_p = math.ceil(64 * 2.0 * kfar * dop * Dp) / 64
        if pulse is not None:
            Dp, pass_p = int(pulse[n][0]), float(pulse[n][1])
        ak, ap = min(pass_k, 0.95), min(pass_p, 0.95)
        mk, Ko = _margin(K, Dk, ak, atten)
        mp, Po = _margin(P, Dp, ap, atten)
        # an axis shorter than its decimation kernel is not decimated
        if Dk > 1 and K < 2 * mk * Dk + Dk:
            Dk, pass_k, ak = 1, rk / (C / (2.0 * df) / 2.0), 0.95
            mk, Ko = 0, K
        if Dp > 1 and P < 2 * mp * Dp + Dp:
            Dp, pass_p, ap = 1, 2.0 * kfar * dop, 0.95
            mp, Po = 0, P
        pidx = Dp * (np.arange(Po) - mp) + (Dp - 1) / 2.0
        out['levels'].append(dict(
            sx=sx, sy=sy, C=sx * sy, Dk=Dk, Dp=Dp, K=K, P=P, Ko=Ko, Po=Po, mk=mk, mp=mp,
            fir_k=decimator_fir(K, Dk, ak, atten)[0] if Dk > 1 else None,
            fir_p=decimator_fir(P, Dp, ap, atten)[0] if Dp > 1 else None,
            args_k=(K, Dk, ak, float(atten)), args_p=(P, Dp, ap, float(atten)),
            f0=f0, df=df, ref=ref, d=cen[0] - ref[0], cen=cen, pidx=pidx,
            pass_k=float(pass_k), pass_p=float(pass_p)))
        ant = positions(ant, pidx)
        f0, df, K, P = f0 + ((Dk - 1) / 2.0 - mk * Dk) * df, df * Dk, Ko, Po
        ref, i0 = cen.reshape(-1, 3), ci0.reshape(-1, 2)
    if (cx, cy) != (T, T):
        raise AssertionError('tile recursion did not end on T x T tiles')
    out['final'] = dict(cen=ref, f0=f0, df=df, K=K, P=P, fc=f0 + (K - 1) / 2.0 * df)
    return out


def pulse_plan(plan):
    """[(Dp, pass_p)] per level and the product of the pulse decimations. (fastsar/ffbp2.py:253-256)"""
    return [(lv['Dp'], lv['pass_p']) for lv in plan['levels']], _prod(lv['Dp'] for lv in plan['levels'])


def plan_signature(plan):
    """The plan's shapes, filters and frequencies as a hashable tuple: equal signatures share compiled kernels and
    filter tables. (fastsar/ffbp2.py:259-273)"""
    def r(x):
        return float(np.float64(x))
    lv = []
    for l in plan['levels']:
        fk = l['fir_k'] and (l['fir_k']['L'], l['fir_k']['pl'], l['fir_k']['pr'], l['fir_k']['n_out'])
        fp = l['fir_p'] and (l['fir_p']['L'], l['fir_p']['pl'], l['fir_p']['pr'], l['fir_p']['n_out'])
        lv.append((l['sx'], l['sy'], l['C'], l['Dk'], l['Dp'], l['K'], l['P'], l['Ko'], l['Po'], r(l['f0']), r(l['df']),
                   fk, fp, l['args_k'], l['args_p'], r(l['pass_k']), r(l['pass_p'])))
    f = plan['final']
    return (tuple(lv), (f['P'], f['K'], r(f['f0']), r(f['df']), r(f['fc'])), plan['T'], plan['nx'], plan['ny'],
            plan['Nx'], plan['Ny'], plan['ox'], plan['oy'], r(plan['spx']), r(plan['spy']),
            tuple(map(r, plan['e1'])), tuple(map(r, plan['e2'])))


def final_phase_error(ant, fmax, nx, ny, spx, spy, e1, e2, T):
    """Predicted image error (dB) of the final stage's model on T x T tiles: the largest phase the model leaves
    out (exact range offset of a pixel from its tile centre minus the linear term and the aperture-mean
    curvature) over a 3 x 3 sample of tile positions and 5 x 5 points of each tile, mapped by
    20 log10(phase) - 14. (fastsar/ffbp2.py:276-297)"""
    ant, e1, e2 = np.asarray(ant, np.float64), np.asarray(e1, np.float64), np.asarray(e2, np.float64)
    a = ant[::max(1, len(ant) // 2048)]
    hx, hy = 0.5 * T * spx, 0.5 * T * spy
    worst = 1e-30
    for fx in (-0.5, 0.0, 0.5):
        for fy in (-0.5, 0.0, 0.5):
            c = fx * max(0.0, nx * spx - 2 * hx) * e1 + fy * max(0.0, ny * spy - 2 * hy) * e2
            w = a - c
            r = np.linalg.norm(w, axis=1)
            u = w / r[:, None]
            for sx in np.linspace(-1, 1, 5):
                for sy in np.linspace(-1, 1, 5):
                    d = sx * hx * e1 + sy * hy * e2
                    ud = u @ d
                    res = np.linalg.norm(w - d, axis=1) - r + ud - (d @ d - ud * ud).mean() / (2 * r.mean())
                    worst = max(worst, float(np.abs(res - res.mean()).max()))
    return 20.0 * np.log10(4 * np.pi * fmax / C * worst) - 14.0


def choose_T(ant, fmax, nx, ny, spx, spy, e1, e2, target_db=-40.0, choices=(32, 16)):
    """The largest tile of `choices` whose predicted error meets target_db, else the last, with its prediction.
    (fastsar/ffbp2.py:300-307)"""
    err = None
    for T in choices:
        err = final_phase_error(ant, fmax, nx, ny, spx, spy, e1, e2, T)
        if err <= target_db:
            return T, err
    return choices[-1], err


def collection_arrays(plan, ant, ref=None):
    """Per-collection float64 work: each level's antenna track as unit vectors u and ranges r0, and for the first
    HOST_LEVELS levels the rotation of every child at every pulse, as the band-centre phase c0 (cycles, wrapped to
    [-1/2, 1/2]) and the slope in cycles per frequency sample, both stored as float32 [B, C, P]:
        ddr = |cen_c - a_p| - base_p,   c0 = wrap(2 (f0 + kc df) / c ddr),   slope = 2 df / c ddr,
    base_p = |ref_parent - a_p| (level 0: |a_p|, or the given per-pulse reference range).
    (fastsar/ffbp2.py:313-342)"""
    ant = np.asarray(ant, np.float64)
    out = dict(levels=[])
    for i, lv in enumerate(plan['levels']):
        r0 = np.linalg.norm(ant, axis=1)
        e = dict(u=ant / r0[:, None], r0=r0)
        if i < HOST_LEVELS:
            kc = (lv['K'] - 1) / 2.0
            B, Cn = lv['ref'].shape[0], lv['C']
            c0 = np.empty((B, Cn, lv['P']), np.float32)
            sl = np.empty((B, Cn, lv['P']), np.float32)
            k0 = 2.0 * (lv['f0'] + kc * lv['df']) / C
            k1 = 2.0 * lv['df'] / C
            for b in range(B):
                base = np.linalg.norm(lv['ref'][b][None, :] - ant, axis=1)
                if i == 0 and ref is not None:
                    base = np.asarray(ref, np.float64)
                ddr = np.linalg.norm(lv['cen'][b][:, None, :] - ant[None, :, :], axis=2) - base[None, :]
                cyc = k0 * ddr
                c0[b] = cyc - np.round(cyc)
                sl[b] = k1 * ddr
            e['c0'], e['slope'] = c0, sl
        out['levels'].append(e)
        ant = positions(ant, lv['pidx'])
    r0 = np.linalg.norm(ant, axis=1)
    out['final'] = dict(u=ant / r0[:, None], r0=r0)
    return out


def level_phase_constants(lv):
    """(k0c, k1) of a level: cycles per metre of range difference at band centre and per frequency sample."""
    return 2.0 * (lv['f0'] + (lv['K'] - 1) / 2.0 * lv['df']) / C, 2.0 * lv['df'] / C


def device_phases(u, r0, d, refs, k0c, k1):
    """The rotation of every child of every parent at a level >= HOST_LEVELS, float64 in, float32 out, in the
    cancellation-free form (no difference of two large ranges):
        wn = |ref - a| = r0 sqrt(1 + (|ref|^2 - 2 r0 u.ref) / r0^2),  wd = r0 (d.u) - d.ref,
        num = |d|^2 - 2 wd,  ddr = |ref + d - a| - |ref - a| = num / (sqrt(wn^2 + num) + wn).
    u [P, 3], r0 [P], d [C, 3], refs [Np, 3] -> c0, slope [Np, C, P]. (fastsar/ffbp2.py:578-590)"""
    uc = refs @ u.T
    wn = r0[None, :] * np.sqrt(1 + ((refs * refs).sum(1)[:, None] - 2 * r0[None, :] * uc) / (r0 * r0)[None, :])
    ud = d @ u.T
    wd = r0[None, None, :] * ud[None] - (d[None, :, :] * refs[:, None, :]).sum(2)[:, :, None]
    num = (d * d).sum(1)[None, :, None] - 2 * wd
    ddr = num / (np.sqrt(wn[:, None, :] ** 2 + num) + wn[:, None, :])
    c0 = ddr * k0c
    return (c0 - np.rint(c0)).astype(np.float32), (ddr * k1).astype(np.float32)


def tile_digits(plan):
    """For the final tiles in plan order (level-0 child major, then level-1 child, ...), the first pixel (X, Y) of
    each tile on the padded Nx x Ny grid: [B, 2] int64. The tile order is that of the levels' children,
    child c = ix * sy + iy at every level. (fastsar/ffbp_cpu.py:172-185)"""
    out = np.zeros((1, 2), np.int64)
    cx, cy = plan['Nx'], plan['Ny']
    for lv in plan['levels']:
        cx, cy = cx // lv['sx'], cy // lv['sy']
        gi, gj = np.meshgrid(np.arange(lv['sx']) * cx, np.arange(lv['sy']) * cy, indexing='ij')
        off = np.stack([gi.ravel(), gj.ravel()], 1)
        out = (out[:, None, :] + off[None, :, :]).reshape(-1, 2)
    return out
