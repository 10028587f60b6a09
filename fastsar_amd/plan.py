"""Host-side planner of factorized backprojection, in float64 NumPy.

Factorized backprojection forms an image by dividing its grid recursively into tiles. A level takes each parent
tile's data, held as [pulses x frequency samples], re-references it in phase to the centre of each child tile, and
low-pass filters and decimates it along both axes; a child needs fewer samples than its parent because it is
smaller (a shorter range extent: coarser frequency sampling suffices) and because the range to it changes less from
pulse to pulse (a subaperture of pulses can be merged). After the last level every T x T tile holds a short
[Pf x Qf] block, which the final stage turns into pixels with a plane-wave model plus a per-pixel quadratic phase.

What this module decides, from the collection geometry and the grid alone:

  choose_splits      how each axis is divided at each level                     (fastsar/ffbp2.py:34-62)
  decimator_fir      the decimation filter of one axis of one level             (fastsar/ffbp2.py:103-137)
  positions          the antenna track on a level's fractional pulse grid      (fastsar/ffbp.py:122-125)
  make_plan          child tiles, decimations and filters of every level        (fastsar/ffbp2.py:192-250)
  final_phase_error  the error the final stage's tile model predicts           (fastsar/ffbp2.py:276-297)
  choose_T           the largest final tile that meets an error target          (fastsar/ffbp2.py:300-307)
  collection_arrays  the rotations of the host levels and every level's track (fastsar/ffbp2.py:313-342)
  pulse_plan, plan_signature  the pulse layout of a plan, and its hashable key  (fastsar/ffbp2.py:253-273)

The plan is a dict with the reference's keys and meanings, so that the images of every backend agree; the
decisions (splits, decimations, margins, filter lengths and placements, output shapes) are the reference's, pinned
by tests/golden/plans.json. The code is this module's own: the decimator is designed from its polyphase form (each
tap divided by the sum of its polyphase branch) and the error model is evaluated on all of a tile's sample points at
once.
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


# ------------------------------------------------------------------------------------------------ decimation filters

class Kaiser:
    """Kaiser's windowed-sinc design of a decimation by D: passband edge `pass_frac` of the output Nyquist rate,
    stopband from 2 - pass_frac, `atten` dB down. Textbook formulas (J. F. Kaiser, 1974): transition width
    dw = 2 pi (1 - pass_frac) / D rad per input sample; window half length (A - 8) / (2.285 dw) / 2 rounded up to a
    half sample, plus one; beta = 0.1102 (A - 8.7) for A > 50 dB. An output reaches `margin` output steps past each
    end of the input."""

    def __init__(self, D, pass_frac, atten):
        self.D, self.pass_frac, self.atten = int(D), float(pass_frac), float(atten)
        dw = (2.0 - 2.0 * self.pass_frac) * np.pi / self.D
        self.half = math.ceil((self.atten - 8.0) / (2.285 * dw)) / 2.0 + 1.0
        self.beta = 0.1102 * (self.atten - 8.7)
        self.margin = int(math.ceil(self.half / self.D))

    def outputs(self, n_in):
        return -(-int(n_in) // self.D) + 2 * self.margin

    def weight(self, d):
        """The windowed sinc at offsets d (input samples) from an output's centre; zero beyond the half length."""
        d = np.asarray(d, np.float64)
        inside = np.abs(d) <= self.half
        arg = np.sqrt(np.clip(1.0 - (d / self.half) ** 2, 0.0, 1.0))
        return np.where(inside, np.sinc(d / self.D) * (np.i0(self.beta * arg) / np.i0(self.beta)), 0.0)


@functools.lru_cache(maxsize=512)
def _polyphase_design(n_in, D, pass_frac, atten):
    """The kernel of decimator_fir, memoized (the design evaluates Bessel functions; patches of a mosaic share it)."""
    k = Kaiser(D, pass_frac, atten)
    m, n_out = k.margin, k.outputs(n_in)
    s = (D - 1) / 2.0                      # output j sits at input D (j - m) + s: offsets from it lie on Z - s
    t_lo = math.ceil(-k.half + s) - s
    t_hi = math.floor(k.half + s) - s
    offs = np.arange(int(round(t_hi - t_lo)) + 1, dtype=np.float64) + t_lo
    # The decimation matrix is normalized so that each input's weights over all outputs sum to one. The outputs
    # that reach an input at offset t from one output sit at t - q D from it (q integer), so its total weight is
    # the sum of the window over that polyphase branch.
    reach = int(math.ceil(2.0 * k.half / D)) + 1           # |t - q D| <= half with |t| <= half: |q| <= 2 half / D
    branch = offs[:, None] - D * np.arange(-reach, reach + 1, dtype=np.float64)[None, :]
    total = k.weight(branch).sum(1)
    kern = k.weight(offs) / total
    L = len(kern)
    # the central output must see the whole kernel inside the input (else the axis is too short to decimate)
    centre = D * (n_out // 2 - m) + s
    if centre + t_lo < 0 or centre + t_hi > n_in - 1:
        raise ValueError(f'an axis of {n_in} samples is shorter than its decimation kernel ({L} taps at D = {D})')
    pl = int(round(D * m - t_lo - s))
    pr = max(D * (n_out - 1) + L - pl - n_in, 0)
    kern.flags.writeable = False
    return kern, pl, pr, n_out, L, m


def decimator_fir(n_in, D, pass_frac, atten=70.0):
    """Kaiser-windowed sinc decimation of n_in samples by D as one shift-invariant kernel:
    out[j] = sum_{r < L} kern[r] * x[D j + r - pl], x zero outside [0, n_in), for j < n_out = ceil(n_in / D) + 2 m.
    Output j is centred on input D (j - m) + (D - 1) / 2; pass_frac is the passband edge as a fraction of the
    output Nyquist rate, the stopband starts at 2 - pass_frac; every input's weights over the outputs sum to one.
    Returns (dict(kern, pl, pr, n_out, L), m). (fastsar/ffbp2.py:103-137, fastsar/ffbp.py:57-78)"""
    kern, pl, pr, n_out, L, m = _polyphase_design(int(n_in), int(D), float(pass_frac), float(atten))
    return dict(kern=kern, pl=pl, pr=pr, n_out=n_out, L=L), m


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


def positions(ant, idx):
    """Antenna positions at fractional pulse indices idx: linear interpolation between neighbouring pulses, and
    linear extrapolation beyond either end. (fastsar/ffbp.py:122-125)"""
    ant = np.asarray(ant, np.float64)
    i0 = np.clip(np.floor(idx).astype(np.int64), 0, len(ant) - 2)
    return ant[i0] + (np.asarray(idx, np.float64) - i0)[:, None] * (ant[i0 + 1] - ant[i0])


# ------------------------------------------------------------------------------------------------ the plan

def _children(first, cx, cy, sx, sy):
    """First pixels [B, sx sy, 2] of the children of parents whose first pixels are first [B, 2]; child c = ix sy + iy
    sits at (ix cx, iy cy) from its parent's first pixel."""
    ix, iy = np.meshgrid(np.arange(sx) * cx, np.arange(sy) * cy, indexing='ij')
    return first[:, None, :] + np.stack([ix.ravel(), iy.ravel()], axis=1)[None, :, :]


def _track_rates(ant, e1, e2, hx, hy):
    """(rk, dop): the largest range offset of a point of a child tile of half extents hx, hy from its centre, over
    the track (metres), and the largest change of that offset from one pulse to the next."""
    u = ant / np.linalg.norm(ant, axis=1)[:, None]
    a1, a2 = u @ e1, u @ e2
    rk = hx * np.abs(a1).max() + hy * np.abs(a2).max()
    dop = hx * np.abs(np.diff(a1)).max() + hy * np.abs(np.diff(a2)).max()
    return rk, dop


def make_plan(col, nx, ny, spx, spy, T=32, nlev=3, pmax=0.4, atten=70.0, splits=None, e1=(1.0, 0.0, 0.0),
              e2=(0.0, 1.0, 0.0), pulse=None):
    """Tile grid, decimation factors and filters of every level for an nx x ny image of spacings spx, spy along the
    orthonormal axes e1, e2, centred on the origin of the collection's coordinates. col: anything with fmin, df,
    K, ant [P, 3] (sim.Collect). pulse: [(Dp, pass_p)] per level in place of the ones this track gives.
    (fastsar/ffbp2.py:192-250)

    A child may keep the fraction pmax of the band its samples resolve: in range, a child of half range extent rk
    needs samples no finer than 2 rk over the unambiguous range c / (2 df) allows, so Dk = floor(pmax c / (4 df rk));
    in pulses, the child's range changes by at most dop per pulse, a phase rate of 2 (2 f_max / c) dop cycles per
    pulse, so Dp = floor(pmax / (2 (2 f_max / c) dop)). Pass bands are the used fractions rounded up to 1/64 of the
    output Nyquist rate (so that near-equal geometries share filters); an axis shorter than its kernel is not
    decimated. A level's output pulse j is centred on its input pulse Dp (j - mp) + (Dp - 1) / 2, and the next
    level's track is the previous one at those fractional indices."""
    e1, e2 = np.asarray(e1, np.float64), np.asarray(e2, np.float64)
    sxs = choose_splits(nx, T, nlev) if splits is None else tuple(int(s[0]) for s in splits)
    sys_ = choose_splits(ny, T, nlev) if splits is None else tuple(int(s[1]) for s in splits)
    Nx, Ny = T * _prod(sxs), T * _prod(sys_)
    ox, oy = (Nx - nx) // 2, (Ny - ny) // 2
    track = np.asarray(col.ant, np.float64)
    f0, df, K, P = float(col.fmin), float(col.df), int(col.K), int(track.shape[0])
    parents_first = np.zeros((1, 2))
    parents_cen = np.zeros((1, 3))
    cx, cy = Nx, Ny
    out = dict(levels=[], T=T, nx=nx, ny=ny, Nx=Nx, Ny=Ny, ox=ox, oy=oy, spx=spx, spy=spy,
               split=list(zip(sxs, sys_)), e1=e1, e2=e2)
    for n, (sx, sy) in enumerate(zip(sxs, sys_)):
        cx, cy = cx // sx, cy // sy
        first = _children(parents_first, cx, cy, sx, sy)                                        # [B, C, 2]
        a = (first[..., 0] + (cx - 1) / 2.0 - ox - nx / 2.0) * spx
        b = (first[..., 1] + (cy - 1) / 2.0 - oy - ny / 2.0) * spy
        cen = a[..., None] * e1 + b[..., None] * e2                                            # [B, C, 3]
        rk, dop = _track_rates(track, e1, e2, 0.5 * cx * spx, 0.5 * cy * spy)
        unamb = C / (2.0 * df)                     # unambiguous range of the level's input samples, m
        kfar = 2.0 * (f0 + K * df) / C             # cycles per metre of two-way range at the top of the band
        Dk = max(1, int(pmax * (unamb / 2.0) / rk))
        Dp = max(1, int(pmax / (2.0 * kfar * dop)))
        pass_k = math.ceil(64 * rk / (C / (2.0 * df * Dk) / 2.0)) / 64
        pass_p = math.ceil(64 * 2.0 * kfar * dop * Dp) / 64
        if pulse is not None:
            Dp, pass_p = int(pulse[n][0]), float(pulse[n][1])
        ak, ap = min(pass_k, 0.95), min(pass_p, 0.95)
        mk = Kaiser(Dk, ak, atten).margin if Dk > 1 else 0
        mp = Kaiser(Dp, ap, atten).margin if Dp > 1 else 0
        if Dk > 1 and K < (2 * mk + 1) * Dk:
            Dk, pass_k, ak, mk = 1, rk / (unamb / 2.0), 0.95, 0
        if Dp > 1 and P < (2 * mp + 1) * Dp:
            Dp, pass_p, ap, mp = 1, 2.0 * kfar * dop, 0.95, 0
        Ko = K if Dk == 1 else Kaiser(Dk, ak, atten).outputs(K)
        Po = P if Dp == 1 else Kaiser(Dp, ap, atten).outputs(P)
        pidx = Dp * (np.arange(Po) - mp) + (Dp - 1) / 2.0
        out['levels'].append(dict(
            sx=sx, sy=sy, C=sx * sy, Dk=Dk, Dp=Dp, K=K, P=P, Ko=Ko, Po=Po, mk=mk, mp=mp,
            fir_k=decimator_fir(K, Dk, ak, atten)[0] if Dk > 1 else None,
            fir_p=decimator_fir(P, Dp, ap, atten)[0] if Dp > 1 else None,
            args_k=(K, Dk, ak, float(atten)), args_p=(P, Dp, ap, float(atten)),
            f0=f0, df=df, ref=parents_cen, d=cen[0] - parents_cen[0], cen=cen, pidx=pidx,
            pass_k=float(pass_k), pass_p=float(pass_p)))
        # the next level: the track on this level's output pulses, the band of its output samples
        track = positions(track, pidx)
        f0 = f0 + ((Dk - 1) / 2.0 - mk * Dk) * df
        df, K, P = df * Dk, Ko, Po
        parents_first, parents_cen = first.reshape(-1, 2), cen.reshape(-1, 3)
    if (cx, cy) != (T, T):
        raise AssertionError('tile recursion did not end on T x T tiles')
    out['final'] = dict(cen=parents_cen, f0=f0, df=df, K=K, P=P, fc=f0 + (K - 1) / 2.0 * df)
    return out


def pulse_plan(plan):
    """A plan's pulse layout: [(Dp, pass_p)] per level and the input pulses one final subaperture spans (the product
    of the decimations); make_plan(pulse=...) lays another collection out the same way. (fastsar/ffbp2.py:253-256)"""
    layout = [(int(lv['Dp']), float(lv['pass_p'])) for lv in plan['levels']]
    return layout, _prod(d for d, _ in layout)


def plan_signature(plan):
    """Everything of a plan that fixes the former's device shapes and filters, as a hashable tuple of integers and
    float64 values: plans with equal signatures can share one former's constants and buffers (a mosaic's patches of
    equal shape and band), with their own geometry. (fastsar/ffbp2.py:259-273)"""
    def filt(fr):
        return None if fr is None else (fr['L'], fr['pl'], fr['pr'], fr['n_out'], fr['kern'].tobytes())
    levels = tuple((lv['sx'], lv['sy'], lv['Dk'], lv['Dp'], lv['K'], lv['P'], lv['Ko'], lv['Po'], float(lv['f0']),
                    float(lv['df']), filt(lv['fir_k']), filt(lv['fir_p'])) for lv in plan['levels'])
    fin = plan['final']
    grid = (plan['T'], plan['nx'], plan['ny'], plan['Nx'], plan['Ny'], plan['ox'], plan['oy'], float(plan['spx']),
            float(plan['spy'])) + tuple(float(v) for v in np.concatenate([plan['e1'], plan['e2']]))
    return levels, (fin['P'], fin['K'], float(fin['f0']), float(fin['df'])), grid


# ------------------------------------------------------------------------------------------------ final-stage error

def final_phase_error(ant, fmax, nx, ny, spx, spy, e1, e2, T):
    """Predicted image error (dB) of the final stage's model on T x T tiles.

    The final stage replaces the range from a pulse's antenna position a to a pixel at offset d from its tile's
    centre c by |w| - u.d plus one curvature term, (|d|^2 - (u.d)^2) / (2 r) averaged over the aperture (w = a - c,
    r = |w|, u = w / r). What it leaves out, |w - d| - r + u.d - that term, less its mean over the pulses (a
    constant phase), is evaluated at 5 x 5 points of a tile for tiles at the centre, edges and corners of the grid
    (3 x 3 positions), on the track thinned to at most about 2048 pulses; its largest value as a phase at f_max is
    mapped by 20 log10(phase) - 14, which follows measured errors to about 2 dB. (fastsar/ffbp2.py:276-297)"""
    ant, e1, e2 = np.asarray(ant, np.float64), np.asarray(e1, np.float64), np.asarray(e2, np.float64)
    a = ant[::max(1, len(ant) // 2048)]
    hx, hy = 0.5 * T * spx, 0.5 * T * spy
    g = np.linspace(-1.0, 1.0, 5)
    gx, gy = np.meshgrid(g * hx, g * hy, indexing='ij')
    d = gx.reshape(-1, 1) * e1 + gy.reshape(-1, 1) * e2                     # [25, 3] points of a tile
    dd = (d * d).sum(1)
    worst = 1e-30
    for fx in (-0.5, 0.0, 0.5):
        for fy in (-0.5, 0.0, 0.5):
            c = fx * max(0.0, nx * spx - 2 * hx) * e1 + fy * max(0.0, ny * spy - 2 * hy) * e2
            w = a - c                                                       # [P, 3]
            r = np.linalg.norm(w, axis=1)
            u = w / r[:, None]
            ud = d @ u.T                                                    # [25, P]
            curv = (dd[:, None] - ud * ud).mean(1) / (2 * r.mean())        # [25]
            exact = np.linalg.norm(w[None, :, :] - d[:, None, :], axis=2)  # [25, P]
            left = exact - r[None, :] + ud - curv[:, None]
            worst = max(worst, float(np.abs(left - left.mean(1, keepdims=True)).max()))
    return 20.0 * np.log10(4 * np.pi * fmax / C * worst) - 14.0


def choose_T(ant, fmax, nx, ny, spx, spy, e1, e2, target_db=-40.0, choices=(32, 16)):
    """The first (largest) final tile of `choices` whose predicted error meets target_db, with its prediction; the
    last choice and its prediction when none does. (fastsar/ffbp2.py:300-307)"""
    pred = None
    for T in choices:
        pred = final_phase_error(ant, fmax, nx, ny, spx, spy, e1, e2, T)
        if pred <= target_db:
            break
    return T, pred


# ------------------------------------------------------------------------------------------------ collection arrays

def level_phase_constants(lv):
    """(k0c, k1) of a level: cycles per metre of range difference at band centre and per frequency sample."""
    return 2.0 * (lv['f0'] + (lv['K'] - 1) / 2.0 * lv['df']) / C, 2.0 * lv['df'] / C


def _host_rotations(lv, track, ref_range=None, block=8):
    """c0, slope [B, C, P] float32 of one host level: the range difference ddr = |cen_c - a_p| - base_p of each child
    centre from its parent's reference (base_p = |ref_parent - a_p|, or the given per-pulse range), as the wrapped
    band-centre phase and the phase per frequency sample. Parents in blocks (bounded temporaries)."""
    k0, k1 = level_phase_constants(lv)
    B, Cn, P = lv['ref'].shape[0], lv['C'], lv['P']
    c0 = np.empty((B, Cn, P), np.float32)
    sl = np.empty((B, Cn, P), np.float32)
    for b0 in range(0, B, block):
        b1 = min(B, b0 + block)
        if ref_range is None:
            base = np.linalg.norm(lv['ref'][b0:b1, None, :] - track[None, :, :], axis=2)           # [nb, P]
        else:
            base = np.broadcast_to(np.asarray(ref_range, np.float64)[None, :], (b1 - b0, P))
        dist = np.linalg.norm(lv['cen'][b0:b1, :, None, :] - track[None, None, :, :], axis=3)   # [nb, C, P]
        ddr = dist - base[:, None, :]
        cyc = k0 * ddr
        c0[b0:b1] = cyc - np.rint(cyc)
        sl[b0:b1] = k1 * ddr
    return c0, sl


def collection_arrays(plan, ant, ref=None):
    """Per-collection float64 work: each level's antenna track as unit vectors u and ranges r0, and for the first
    HOST_LEVELS levels the rotation of every child at every pulse (c0 in cycles, wrapped to [-1/2, 1/2], and the
    slope in cycles per frequency sample, float32 [B, C, P]). ref [P]: each pulse's reference range, when it is not
    the antenna's distance to the origin; the first level's rotations start from it. (fastsar/ffbp2.py:313-342)"""
    track = np.asarray(ant, np.float64)
    levels = []
    for n, lv in enumerate(plan['levels']):
        r0 = np.linalg.norm(track, axis=1)
        e = dict(u=track / r0[:, None], r0=r0)
        if n < HOST_LEVELS:
            e['c0'], e['slope'] = _host_rotations(lv, track, ref if n == 0 else None)
        levels.append(e)
        track = positions(track, lv['pidx'])
    r0 = np.linalg.norm(track, axis=1)
    return dict(levels=levels, final=dict(u=track / r0[:, None], r0=r0))


def device_phases(u, r0, d, refs, k0c, k1):
    """The rotation of every child of every parent at a level >= HOST_LEVELS, float64 in, float32 out, with no
    difference of two large ranges: for the parent reference q and the child offset d,
        ddr = |q + d - a| - |q - a| = (|d|^2 + 2 d.(q - a)) / (|q + d - a| + |q - a|),
    each norm formed from the vector difference in float64. u [P, 3], r0 [P] (a = r0 u), d [C, 3], refs [Np, 3]
    -> c0, slope [Np, C, P]. (fastsar/ffbp2.py:578-590)"""
    a = u * r0[:, None]                                                      # [P, 3]
    qa = refs[:, None, :] - a[None, :, :]                                    # [Np, P, 3]
    near = np.linalg.norm(qa, axis=2)                                        # [Np, P]
    far = np.linalg.norm(qa[:, None, :, :] + d[None, :, None, :], axis=3)   # [Np, C, P]
    num = (d * d).sum(1)[None, :, None] + 2 * np.einsum('ck,npk->ncp', d, qa)
    ddr = num / (far + near[:, None, :])
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
        out = _children(out, cx, cy, lv['sx'], lv['sy']).reshape(-1, 2).astype(np.int64)
    return out
