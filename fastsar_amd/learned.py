"""Learned decimation filters for the factorized backprojection's levels.

Each level decimates its children along range (samples) and along the aperture (pulses) with a polyphase filter:
output j is sum_r h[r] x[D j + r - pl], tap r = D u + ph belonging to phase ph (the input's fractional offset
ph / D within an output period) and tap u of that phase. The plan's filters are Kaiser-windowed sincs
(fastsar/ffbp2.py:103-137), long enough for the design attenuation. Here a small network emits T taps per phase
instead (L = D T taps in all), trained so that the level's outputs approach those of an exact decimator:

  * network: an MLP [f, s] -> tanh 64 -> tanh 64 -> T per axis (one for range, one for pulses), f = (ph + 1/2) / D
    the phase's fractional offset, s = level / (levels - 1) the normalised stage index; its output is added to a
    Kaiser-windowed sinc of the same length (the fallback filter, so an untrained network gives the fallback);
  * pairs: the method itself, run on seeded scenes with its own (Kaiser) filters up to each level: the level's
    input windows (the rotated parent rows for range, the range-decimated rows for pulses) against the same
    inputs decimated by a 140 dB Kaiser decimator (the exact merged samples), both seen through a sharp low-pass
    at the output rate that keeps the band the image is formed from (content the decimation leaves in the
    transition band never reaches a pixel, so it is not fitted); they enter training through their
    least-squares statistics R, v, E (fsh_ls_stats), which give every pair's squared error exactly;
  * training: the relative squared error per stage, averaged, minimised by Adam, with the forward pass, the
    least-squares gradient, the backward pass and the update all hand-written HIP kernels (fsh_tapnet_*).

    st = Stats(T); st.add_scene(plan, coll, S_dev)            # for each training scene
    taps = train(st, iters=3000)                              # LearnedTaps
    plan2 = taps.apply(plan)                                  # the plan with learned filters (same Ko, Po)
"""
import ctypes
import math

import numpy as np

from . import _hip
from . import plan as _plan

HID = 64
MAXL = 128
EXACT_ATTEN = 140.0
PB_M = 12                # half length of the passband filter of the training pairs


def _bind():
    L = _hip.lib()
    if not getattr(L, '_learned_bound', False):
        vp, ci, cl, cf = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_float
        L.fsh_ls_stats.argtypes = [vp, vp, cl, ci, cl, cl, cl, cl, cl, cl, ci, ci, ci, ci, ci, ci, vp, vp, vp, vp]
        L.fsh_tapnet_nparam.argtypes = [ci]
        L.fsh_tapnet_forward.argtypes = [vp, ci, vp, ci, vp, vp, vp]
        L.fsh_tapnet_backward.argtypes = [vp, ci, vp, ci, vp, vp, vp, vp, vp]
        L.fsh_tapnet_ls_grad.argtypes = [vp, ci, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.fsh_adam.argtypes = [vp, vp, vp, vp, ci, cf, cf, cf, cf, ci, vp]
        for n in ('fsh_ls_stats', 'fsh_tapnet_nparam', 'fsh_tapnet_forward', 'fsh_tapnet_backward', 'fsh_tapnet_ls_grad',
                  'fsh_adam'):
            getattr(L, n).restype = ci
        L._learned_bound = True
    return L


def nparam(T):
    return 4 * HID + HID * HID + T * HID + T


# ------------------------------------------------------------------------------------------------ filters

def placement(D, m, T):
    """(L, pl) of a T-taps-per-phase filter for a decimation by D whose output j is centred on input
    D (j - m) + (D - 1) / 2 (the plan's convention): L = D T, and pl puts the window's centre on the output's
    (half a sample early when D (T - 1) is odd)."""
    L = D * T
    return L, D * m + (D * (T - 1)) // 2


def kaiser_same_length(D, T, pass_frac, m):
    """The fallback filter at the learned length: a Kaiser-windowed sinc of L = D T taps about the output centre,
    its attenuation the most Kaiser's length formula allows for the transition (2 - 2 pass) pi / D, each phase's
    taps summing to one (the plan's row-sum normalisation). -> float64 [L]."""
    L, pl = placement(D, m, T)
    dw = (2.0 - 2.0 * pass_frac) * np.pi / D
    half = (L - 1) / 2.0
    A = 8.0 + 2.285 * dw * 2.0 * max(half - 1.0, 0.5)
    if A > 50:
        beta = 0.1102 * (A - 8.7)
    elif A > 21:
        beta = 0.5842 * (A - 21) ** 0.4 + 0.07886 * (A - 21)
    else:
        beta = 0.0
    # tap r sits at input D j + r - pl; the output centre at D j - D m + (D - 1) / 2
    d = np.arange(L) - pl + D * m - (D - 1) / 2.0
    h = np.sinc(d / D) * np.i0(beta * np.sqrt(np.clip(1.0 - (d / (half + 1.0)) ** 2, 0.0, 1.0))) / np.i0(beta)
    for ph in range(D):
        s = h[ph::D].sum()
        if s != 0:
            h[ph::D] /= s
    return h


def passband_filter(pass_frac, M=32):
    """A sharp low-pass at the output rate keeping |f| < pass_frac of the output Nyquist rate (the band the image
    is formed from): a Kaiser-windowed sinc of 2 M + 1 taps, unit DC gain. float32 [2 M + 1]."""
    n = np.arange(-M, M + 1)
    wc = min(pass_frac, 0.95)
    h = wc * np.sinc(wc * n) * np.i0(8.0 * np.sqrt(1.0 - (n / (M + 1.0)) ** 2)) / np.i0(8.0)
    return (h / h.sum()).astype(np.float32)


def exact_filter(n_in, D, pass_frac, m):
    """The exact decimator of a level (Kaiser at 140 dB) placed so that its output j has the plan's centre:
    (taps float32, L, pl)."""
    fr, m_long = _plan.decimator_fir(n_in, D, min(pass_frac, 0.95), EXACT_ATTEN)
    pl = fr['pl'] - D * (m_long - m)
    return np.ascontiguousarray(fr['kern'], np.float32), fr['L'], pl


# ------------------------------------------------------------------------------------------------ statistics

class Stats:
    """Least-squares statistics of every (level, axis) decimation stage over the scenes added, for filters of T taps
    per phase: R [L x L], v [L], E (float64, device), with the stage's D, m, pass band and L, pl."""

    def __init__(self, T):
        if T < 1 or T > 16:
            raise ValueError('T: 1 to 16 taps per phase')
        self.T = T
        self.stages = {}              # (level, axis) -> dict(D, m, pass, L, pl, R, v, E, pairs)
        self.nlev = None

    def _stage(self, key, D, m, pass_frac):
        st = self.stages.get(key)
        if st is None:
            L, pl = placement(D, m, self.T)
            if L > MAXL:
                raise ValueError(f'a decimation by {D} with {self.T} taps per phase needs {L} taps; at most {MAXL}')
            st = dict(D=D, m=m, pass_frac=pass_frac, L=L, pl=pl, pairs=0, edges=False,
                      R=_hip.DeviceArray((L, L), np.float64).zero(), v=_hip.DeviceArray((L,), np.float64).zero(),
                      E=_hip.DeviceArray((1,), np.float64).zero())
            self.stages[key] = st
        elif (st['D'], st['m']) != (D, m):
            raise ValueError(f'stage {key} seen with decimation ({D}, m {m}) and ({st["D"]}, m {st["m"]})')
        return st

    def _accumulate(self, st, x, y, nlines, inner, xq, xr, xs, yq, yr, ys, n_in, n_out, exact):
        """the pairs of the interior outputs: those whose exact window and passband filter see no zero padding
        (at the edges every short filter differs from the long one in ways no pixel sees after the window); every
        output where the axis has fewer than 8 interior ones"""
        L = _bind()
        D, el, epl = st['D'], exact[1], exact[2]
        j0 = -(-epl // D) + PB_M // 2
        j1 = (n_in + epl - el) // D + 1 - PB_M // 2
        if j1 - j0 < 8:                 # an axis too short for an interior (a deep level's few samples): all outputs
            j0, j1 = 0, n_out
            st['edges'] = True
        _hip.check(L.fsh_ls_stats(_hip.ptr(x), _hip.ptr(y), nlines, inner, xq, xr, xs, yq, yr, ys, n_in, j0, j1 - j0, D,
                                  st['pl'], st['L'], _hip.ptr(st['R']), _hip.ptr(st['v']), _hip.ptr(st['E']), None),
                   'ls_stats')
        st['pairs'] += nlines * (j1 - j0)

    def add_scene(self, plan, coll, S, window=None):
        """Run the method (the plan's own filters) on the history S (DeviceArray [P, K] complex64; windowed by
        window = (wp, wk) on the host copy when given) and add every decimation stage's pairs."""
        L = _hip.lib()
        levels = plan['levels']
        nlev = len(levels)
        if self.nlev is None:
            self.nlev = nlev
        elif self.nlev != nlev:
            raise ValueError('all scenes of one training set share the number of levels')
        if window is not None:
            h = S.to_host() * (window[0][:, None] * window[1][None, :]).astype(np.float32)
            S = _hip.DeviceArray.from_host(h.astype(np.complex64))
        G = levels[0]['C']
        one = _hip.DeviceArray.from_host(np.ones(1, np.float32))

        def rot(a, Np, P, K, c0, sl, Cn, taps, Lt, pl, D, Ko):
            y = _hip.DeviceArray((Np * Cn, P, Ko), np.complex64)
            _hip.check(L.fsh_rot_fir_k(_hip.ptr(a), Np, P, K, _hip.ptr(c0), _hip.ptr(sl), Cn, _hip.ptr(taps), Lt, pl, D,
                                       Ko, (K - 1) / 2.0, _hip.ptr(y), None), 'rot_fir_k')
            return y

        def firp(y, B, P, Ko, taps, Lt, pl, D, Po):
            z = _hip.DeviceArray((B, Po, Ko), np.complex64)
            _hip.check(L.fsh_fir_p(_hip.ptr(y), B, P, Ko, _hip.ptr(taps), Lt, pl, D, Po, _hip.ptr(z), None), 'fir_p')
            return z

        for g in range(G):
            a, Np = S, 1
            for i, lv in enumerate(levels):
                cl = coll['levels'][i]
                P, K, Ko, Po, Dk, Dp = lv['P'], lv['K'], lv['Ko'], lv['Po'], lv['Dk'], lv['Dp']
                if i == 0:
                    c0, sl, Cn = cl['c0'][0, g][None, None], cl['slope'][0, g][None, None], 1
                elif i < _plan.HOST_LEVELS:
                    c0, sl, Cn = cl['c0'][g][None], cl['slope'][g][None], lv['C']
                else:
                    refs = lv['ref'].reshape(G, Np, 3)[g]
                    k0c, k1 = _plan.level_phase_constants(lv)
                    c0, sl = _plan.device_phases(cl['u'], cl['r0'], lv['d'], refs, k0c, k1)
                    Cn = lv['C']
                dc0 = _hip.DeviceArray.from_host(np.ascontiguousarray(c0, np.float32))
                dsl = _hip.DeviceArray.from_host(np.ascontiguousarray(sl, np.float32))
                B = Np * Cn
                if Dk > 1:
                    st = self._stage((i, 'range'), Dk, lv['mk'], lv['pass_k'])
                    pb = passband_filter(lv['pass_k'], PB_M)
                    # the rotated rows filtered by the output-rate passband filter spread to the input rate, and the
                    # exact outputs filtered by it: the pairs then measure the error in the band the image uses
                    pu = np.zeros(2 * PB_M * Dk + 1, np.float32)
                    pu[::Dk] = pb
                    z = rot(a, Np, P, K, dc0, dsl, Cn, _hip.DeviceArray.from_host(pu), len(pu), PB_M * Dk, 1, K)
                    te, tl, tp = exact_filter(K, Dk, lv['pass_k'], lv['mk'])
                    ye = rot(a, Np, P, K, dc0, dsl, Cn, _hip.DeviceArray.from_host(te), tl, tp, Dk, Ko)
                    zero = _hip.DeviceArray((B, 1, P), np.float32).zero()
                    yef = rot(ye, B, P, Ko, zero, zero, 1, _hip.DeviceArray.from_host(pb), len(pb), PB_M, 1, Ko)
                    self._accumulate(st, z, yef, B * P, 1, K, 0, 1, Ko, 0, 1, K, Ko, (te, tl, tp))
                    del z, ye, yef
                    fk = lv['fir_k']
                    y = rot(a, Np, P, K, dc0, dsl, Cn, _hip.DeviceArray.from_host(np.asarray(fk['kern'], np.float32)),
                            fk['L'], fk['pl'], Dk, Ko)
                else:
                    y = rot(a, Np, P, K, dc0, dsl, Cn, one, 1, 0, 1, K)
                if Dp > 1:
                    st = self._stage((i, 'pulse'), Dp, lv['mp'], lv['pass_p'])
                    pb = passband_filter(lv['pass_p'], PB_M)
                    pu = np.zeros(2 * PB_M * Dp + 1, np.float32)
                    pu[::Dp] = pb
                    yf = firp(y, B, P, Ko, _hip.DeviceArray.from_host(pu), len(pu), PB_M * Dp, 1, P)
                    te, tl, tp = exact_filter(P, Dp, lv['pass_p'], lv['mp'])
                    ze = firp(y, B, P, Ko, _hip.DeviceArray.from_host(te), tl, tp, Dp, Po)
                    zef = firp(ze, B, Po, Ko, _hip.DeviceArray.from_host(pb), len(pb), PB_M, 1, Po)
                    self._accumulate(st, yf, zef, B * Ko, Ko, P * Ko, 1, Ko, Po * Ko, 1, Ko, P, Po, (te, tl, tp))
                    del yf, ze, zef
                    fp = lv['fir_p']
                    a = firp(y, B, P, Ko, _hip.DeviceArray.from_host(np.asarray(fp['kern'], np.float32)), fp['L'],
                             fp['pl'], Dp, Po)
                else:
                    a = y
                Np = B
        _hip.check(L.fsh_synchronize(), 'synchronize')

    def host(self):
        """{(level, axis): dict(D, m, pass_frac, L, pl, pairs, R, v, E)} with the statistics on the host."""
        out = {}
        for k, st in self.stages.items():
            out[k] = dict(st, R=st['R'].to_host(), v=st['v'].to_host(), E=float(st['E'].to_host()[0]))
        return out


# ------------------------------------------------------------------------------------------------ network

class TapNet:
    """The MLP's parameters (float32, device) with forward and backward passes on the device."""

    def __init__(self, T, seed=0, params=None):
        self.T = T
        self.n = nparam(T)
        if params is None:
            rng = np.random.default_rng(seed)
            W1 = rng.standard_normal((HID, 2)) * 1.0
            b1 = rng.standard_normal(HID) * 0.5
            W2 = rng.standard_normal((HID, HID)) / np.sqrt(HID)
            b2 = np.zeros(HID)
            W3 = np.zeros((T, HID))                      # the untrained network adds nothing to the fallback
            b3 = np.zeros(T)
            params = np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3]).astype(np.float32)
        params = np.ascontiguousarray(params, np.float32)
        if params.shape != (self.n,):
            raise ValueError(f'a network of {T} taps has {self.n} parameters, got {params.shape}')
        self.p = _hip.DeviceArray.from_host(params)

    def params(self):
        return self.p.to_host()

    def forward(self, pts):
        """pts DeviceArray [n, 2] -> (out DeviceArray [n, T], act DeviceArray [n, 128])."""
        L = _bind()
        n = pts.shape[0]
        out = _hip.DeviceArray((n, self.T), np.float32)
        act = _hip.DeviceArray((n, 2 * HID), np.float32)
        _hip.check(L.fsh_tapnet_forward(_hip.ptr(self.p), self.T, _hip.ptr(pts), n, _hip.ptr(act), _hip.ptr(out), None),
                   'tapnet_forward')
        return out, act

    def backward(self, pts, act, dout, grow=None, grad=None):
        """d loss / d params [nparam] (DeviceArray) from d loss / d out [n, T]."""
        L = _bind()
        n = pts.shape[0]
        grow = grow if grow is not None else _hip.DeviceArray((n, self.n), np.float32)
        grad = grad if grad is not None else _hip.DeviceArray((self.n,), np.float32)
        _hip.check(L.fsh_tapnet_backward(_hip.ptr(self.p), self.T, _hip.ptr(pts), n, _hip.ptr(act), _hip.ptr(dout),
                                         _hip.ptr(grow), _hip.ptr(grad), None), 'tapnet_backward')
        return grad


class Problem:
    """The training problem of one axis: its stages' statistics on the device, the points (f, s) of every phase
    of every stage, the fallback taps; loss_grad(net) -> (loss per stage, d loss / d out)."""

    def __init__(self, stats, axis, nlev, T):
        keys = sorted(k for k in stats if k[1] == axis and stats[k]['E'] > 0)
        if not keys:
            raise ValueError(f'no {axis} stages in the statistics')
        self.keys, self.T = keys, T
        pts, D, Ls, off, R_at, Rs, vs, Es, base = [], [], [], [], [], [], [], [], []
        ofs = 0
        for k in keys:
            st = stats[k]
            d = st['D']
            s = k[0] / max(1, nlev - 1)
            off.append(len(pts))
            pts += [((ph + 0.5) / d, s) for ph in range(d)]
            D.append(d)
            Ls.append(st['L'])
            R_at.append(ofs)
            ofs += st['L'] * st['L']
            Rs.append(np.asarray(st['R'], np.float64).ravel())
            vv = np.zeros(MAXL)
            vv[:st['L']] = st['v']
            vs.append(vv)
            Es.append(st['E'])
            bb = np.zeros(MAXL)
            bb[:st['L']] = kaiser_same_length(d, T, st['pass_frac'], st['m'])
            base.append(bb)
        self.host_base = base
        self.stats = stats
        up = _hip.DeviceArray.from_host
        self.pts = up(np.asarray(pts, np.float32))
        self.D, self.L, self.off = up(np.asarray(D, np.int32)), up(np.asarray(Ls, np.int32)), up(np.asarray(off, np.int32))
        self.R_at = up(np.asarray(R_at, np.int64))
        self.R, self.v, self.E = up(np.concatenate(Rs)), up(np.concatenate(vs)), up(np.asarray(Es, np.float64))
        self.base = up(np.concatenate(base))
        self.wt = up(np.ones(len(keys), np.float32) / len(keys))
        self.npts = len(pts)
        self.dout = _hip.DeviceArray((self.npts, T), np.float32)
        self.loss = _hip.DeviceArray((len(keys),), np.float64)

    def loss_grad(self, out):
        L = _bind()
        _hip.check(L.fsh_tapnet_ls_grad(_hip.ptr(out), self.T, len(self.keys), _hip.ptr(self.D), _hip.ptr(self.L),
                                        _hip.ptr(self.off), _hip.ptr(self.R_at), _hip.ptr(self.R), _hip.ptr(self.v),
                                        _hip.ptr(self.E), _hip.ptr(self.base), _hip.ptr(self.wt), _hip.ptr(self.dout),
                                        _hip.ptr(self.loss), None), 'tapnet_ls_grad')
        return self.loss, self.dout

    def relative_error(self, taps_of):
        """{stage: relative squared error} of filters taps_of(key) -> [L] (host float64), from the statistics."""
        out = {}
        for k in self.keys:
            st = self.stats[k]
            h = np.asarray(taps_of(k), np.float64)
            R, v, E = np.asarray(st['R']), np.asarray(st['v']), st['E']
            out[k] = float((h @ R @ h - 2 * h @ v + E) / E)
        return out


class LearnedTaps:
    """Trained networks of both axes with the fallback filters they correct: taps(level, axis, D, m, pass) -> [L]
    float64, apply(plan) -> the plan with learned filters."""

    def __init__(self, T, nlev, nets, history=None, keep=()):
        self.T, self.nlev, self.nets = T, nlev, nets            # nets: {'range': params, 'pulse': params}
        self.history = history or {}
        self.keep = {tuple(k) for k in keep}                     # (level, axis) stages left with the plan's filters

    def taps(self, level, axis, D, m, pass_frac):
        p = self.nets.get(axis)
        base = kaiser_same_length(D, self.T, pass_frac, m)
        if p is None:
            return base
        f = (np.arange(D) + 0.5) / D
        s = level / max(1, self.nlev - 1)
        out = mlp_host(p, self.T, np.stack([f, np.full(D, s)], 1))           # [D, T]
        h = base.copy()
        for ph in range(D):
            h[ph::D] += out[ph, :len(h[ph::D])]
        return h

    def apply(self, plan):
        """A copy of the plan whose decimation filters are the learned ones (same Ko, Po and output centres)."""
        import copy
        out = copy.deepcopy(plan)
        if len(out['levels']) != self.nlev:
            raise ValueError(f'the filters were trained for {self.nlev} levels, the plan has {len(out["levels"])}')
        for i, lv in enumerate(out['levels']):
            if lv['Dk'] > 1 and (i, 'range') not in self.keep:
                h = self.taps(i, 'range', lv['Dk'], lv['mk'], lv['pass_k'])
                L, pl = placement(lv['Dk'], lv['mk'], self.T)
                lv['fir_k'] = dict(kern=h.astype(np.float64), pl=pl, pr=max(lv['Dk'] * (lv['Ko'] - 1) + L - pl - lv['K'], 0),
                                   n_out=lv['Ko'], L=L)
            if lv['Dp'] > 1 and (i, 'pulse') not in self.keep:
                h = self.taps(i, 'pulse', lv['Dp'], lv['mp'], lv['pass_p'])
                L, pl = placement(lv['Dp'], lv['mp'], self.T)
                lv['fir_p'] = dict(kern=h.astype(np.float64), pl=pl, pr=max(lv['Dp'] * (lv['Po'] - 1) + L - pl - lv['P'], 0),
                                   n_out=lv['Po'], L=L)
        return out

    def save(self, path):
        import json
        hist = json.dumps(self.history, default=float)
        keep = np.array([[k[0], 0 if k[1] == 'range' else 1] for k in sorted(self.keep)], np.int64).reshape(-1, 2)
        np.savez(path, T=self.T, nlev=self.nlev, history=np.array(hist), keep=keep,
                 **{f'net_{k}': v for k, v in self.nets.items()})

    @classmethod
    def load(cls, path):
        import json
        z = np.load(path, allow_pickle=False)
        nets = {k[4:]: z[k] for k in z.files if k.startswith('net_')}
        hist = json.loads(str(z['history'])) if 'history' in z.files else {}
        keep = [(int(a), 'range' if b == 0 else 'pulse') for a, b in z['keep']] if 'keep' in z.files else ()
        return cls(int(z['T']), int(z['nlev']), nets, hist, keep)


def mlp_host(params, T, pts):
    """The network's forward pass on the host (float64): pts [n, 2] -> [n, T]."""
    p = np.asarray(params, np.float64)
    W1 = p[:2 * HID].reshape(HID, 2)
    b1 = p[2 * HID:3 * HID]
    W2 = p[3 * HID:3 * HID + HID * HID].reshape(HID, HID)
    b2 = p[3 * HID + HID * HID:4 * HID + HID * HID]
    W3 = p[4 * HID + HID * HID:4 * HID + HID * HID + T * HID].reshape(T, HID)
    b3 = p[4 * HID + HID * HID + T * HID:]
    h1 = np.tanh(np.asarray(pts, np.float64) @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3.T + b3


def train(stats, iters=3000, lr=1e-3, seed=0, log=None):
    """Train one network per axis on the statistics (Stats) with Adam on the device. -> LearnedTaps with
    .history = {axis: dict(fallback, ls_optimum, learned)} of mean relative errors (dB) over the stages."""
    L = _bind()
    T = stats.T
    hs = stats.host()
    # stages whose axis is too short for interior pairs keep the plan's (Kaiser) filters: near the edges of a short
    # axis no filter of T taps per phase matches the long one well enough
    keep = sorted(k for k, st in hs.items() if st['edges'])
    hs = {k: v for k, v in hs.items() if k not in keep}
    nets, hist = {}, {}
    for axis in ('range', 'pulse'):
        if not any(k[1] == axis for k in hs):
            continue
        prob = Problem(hs, axis, stats.nlev, T)
        net = TapNet(T, seed=seed)
        m = _hip.DeviceArray((net.n,), np.float32).zero()
        v = _hip.DeviceArray((net.n,), np.float32).zero()
        grow = _hip.DeviceArray((prob.npts, net.n), np.float32)
        grad = _hip.DeviceArray((net.n,), np.float32)
        for t in range(1, iters + 1):
            out, act = net.forward(prob.pts)
            loss, dout = prob.loss_grad(out)
            net.backward(prob.pts, act, dout, grow, grad)
            rate = lr if t <= 0.7 * iters else lr * 0.1
            _hip.check(L.fsh_adam(_hip.ptr(net.p), _hip.ptr(grad), _hip.ptr(m), _hip.ptr(v), net.n, rate, 0.9, 0.999, 1e-8,
                                  t, None), 'adam')
            if log and (t == 1 or t % 500 == 0):
                log(f'{axis} iteration {t}: mean relative error {10 * np.log10(loss.to_host().mean()):.2f} dB')
        nets[axis] = net.params()
        lt = LearnedTaps(T, stats.nlev, {axis: nets[axis]}, keep=keep)
        err = lambda taps_of: 10 * np.log10(np.mean(list(prob.relative_error(taps_of).values())))
        opt = {k: np.linalg.solve(hs[k]['R'], hs[k]['v']) for k in prob.keys}
        hist[axis] = dict(
            fallback=err(lambda k: kaiser_same_length(hs[k]['D'], T, hs[k]['pass_frac'], hs[k]['m'])),
            ls_optimum=err(lambda k: opt[k]),
            learned=err(lambda k: lt.taps(k[0], axis, hs[k]['D'], hs[k]['m'], hs[k]['pass_frac'])),
            stages=len(prob.keys), pairs=int(sum(hs[k]['pairs'] for k in prob.keys)))
    return LearnedTaps(T, stats.nlev, nets, hist, keep)


# ------------------------------------------------------------------------------------------------ the small spotlight set

def c1_geometry(P=1024):
    """The small spotlight geometry at P pulses: X band, 300 MHz over 1024 samples, 3 degrees at 10 km slant range and
    30 degrees grazing, a 512 x 512 grid at 0.5 m; the plan of three levels at a final tile of 16."""
    from . import sim
    col = sim.spotlight(P, 1024, 3.0)
    grid = dict(nx=512, ny=512, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    plan = _plan.make_plan(col, grid['nx'], grid['ny'], grid['spx'], grid['spy'], T=16, nlev=3, e1=grid['e1'], e2=grid['e2'])
    return col, grid, plan


def c1_scene(col, seed):
    """25 targets with Rayleigh amplitudes uniform over a 256 m square plus white clutter 30 dB below them, simulated
    on the device (seeded)."""
    from . import sim
    from .exact import simulate_device
    rng = np.random.default_rng(seed)
    pos, amp = sim.point_scene(rng, 25, 256.0, 256.0)
    return simulate_device(col, pos, amp, clutter_sigma=sim.clutter_sigma(amp, -30.0), seed=seed)


def train_c1(T, scenes=200, iters=3000, seed0=1000, log=None):
    """Train the filters of T taps per phase on `scenes` seeded scenes of the small spotlight geometry at 1024 pulses.
    -> LearnedTaps (history: the training errors, the pair count)."""
    from .api import _window
    col, grid, plan = c1_geometry()
    coll = _plan.collection_arrays(plan, col.ant)
    win = _window(col.Np, col.K)
    st = Stats(T)
    for i in range(scenes):
        st.add_scene(plan, coll, c1_scene(col, seed0 + i), window=win)
        if log and (i + 1) % 50 == 0:
            log(f'{i + 1} scenes: {sum(s["pairs"] for s in st.stages.values()):.3g} pairs')
    return train(st, iters=iters, log=log)


def train_scenes(plan, coll, scenes, T, window=None, iters=3000, log=None):
    """Train filters of T taps per phase for one plan on device histories (an iterable of DeviceArrays). -> LearnedTaps."""
    st = Stats(T)
    for i, S in enumerate(scenes):
        st.add_scene(plan, coll, S, window=window)
        if log:
            log(f'scene {i + 1}: {sum(s["pairs"] for s in st.stages.values()):.3g} pairs')
    return train(st, iters=iters, log=log)


def default_taps(T=6):
    """The committed filters of T taps per phase trained on 200 scenes of the small spotlight geometry, or None."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data', f'learned_c1_T{T}.npz')
    return LearnedTaps.load(path) if os.path.exists(path) else None


if __name__ == '__main__':
    import argparse
    import os
    import sys
    ap = argparse.ArgumentParser(description='train the learned decimation filters on the GPU')
    ap.add_argument('--T', type=int, default=6)
    ap.add_argument('--scenes', type=int, default=200)
    ap.add_argument('--iters', type=int, default=3000)
    ap.add_argument('--out', default=None)
    a = ap.parse_args()
    lg = lambda *x: print('[learned]', *x, file=sys.stderr, flush=True)
    taps = train_c1(a.T, a.scenes, a.iters, log=lg)
    lg('training errors (dB):', taps.history)
    out = a.out or os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data', f'learned_c1_T{a.T}.npz')
    os.makedirs(os.path.dirname(out), exist_ok=True)
    taps.save(out)
    lg('saved', out)
