"""The HIP backend factory of factorized backprojection: the plan and collection arrays packed into the C ABI's
descriptors and handed to the native former, which keeps them on the device and forms images on its own stream.

    form = make_ffbp_hip(plan, coll, wf=None, window=(wp, wk))
    img = form(S)                # S [P, K] complex64 (host array or DeviceArray) -> complex64 [nx, ny]

The counterpart of make_ffbp_cpu / make_ffbp_cuda (fastsar/ffbp_cpu.py:158-293, fastsar/ffbp_cuda.py:889-1213).
"""
import ctypes

import numpy as np

from . import _hip
from . import memory as _mem
from .plan import C, HOST_LEVELS, level_phase_constants, tile_digits

STAGES = ('prepare', 'rot_fir_k', 'fir_p', 'device_phases', 'final_geom', 'final_tiles', 'level_fused', 'level_mfma')


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


class HipFFBP:
    """The native former of one plan (fsh_ffbp_create); call with a history to form an image."""

    def __init__(self, plan, coll, wf=None, window=None):
        L = _hip.lib()
        levels, T = plan['levels'], plan['T']
        if T not in (4, 8, 16, 32):
            raise ValueError(f'the hip final stage takes T = 4, 8, 16 or 32, not {T}')
        fin = plan['final']
        self.plan = plan
        self.nx, self.ny = plan['nx'], plan['ny']
        self.G = levels[0]['C']
        self._group_req, self._warned_group = 0, None
        self.P, self.K = levels[0]['P'], levels[0]['K']
        keep = []                                   # host arrays referenced by the descriptors until create returns

        def p32(a):
            if a is None:
                return None
            a = _f32(a)
            keep.append(a)
            return a.ctypes.data_as(_hip.P_f32)

        def p64(a):
            if a is None:
                return None
            a = _f64(a)
            keep.append(a)
            return a.ctypes.data_as(_hip.P_f64)

        descs = (_hip.LevelDesc * len(levels))()
        for i, (lv, cl) in enumerate(zip(levels, coll['levels'])):
            d = descs[i]
            d.sx, d.sy, d.C, d.Dk, d.Dp = lv['sx'], lv['sy'], lv['C'], lv['Dk'], lv['Dp']
            d.K, d.P, d.Ko, d.Po = lv['K'], lv['P'], lv['Ko'], lv['Po']
            d.nparents = lv['ref'].shape[0]
            d.kc = (lv['K'] - 1) / 2.0
            d.k0c, d.k1 = level_phase_constants(lv)
            if lv['Dk'] > 1:
                d.Lk, d.plk, d.taps_k = lv['fir_k']['L'], lv['fir_k']['pl'], p32(lv['fir_k']['kern'])
            if lv['Dp'] > 1:
                d.Lp, d.plp, d.taps_p = lv['fir_p']['L'], lv['fir_p']['pl'], p32(lv['fir_p']['kern'])
            if i < HOST_LEVELS:
                d.c0, d.sl = p32(cl['c0']), p32(cl['slope'])
            else:
                d.u, d.r0, d.d, d.ref = p64(cl['u']), p64(cl['r0']), p64(lv['d']), p64(lv['ref'])
        pd = _hip.PlanDesc()
        pd.nlev, pd.T, pd.nx, pd.ny = len(levels), T, plan['nx'], plan['ny']
        pd.Nx, pd.Ny, pd.ox, pd.oy = plan['Nx'], plan['Ny'], plan['ox'], plan['oy']
        pd.Pf, pd.Qf, pd.P, pd.K = fin['P'], fin['K'], self.P, self.K
        pd.ntiles = fin['cen'].shape[0]
        pd.spx, pd.spy = plan['spx'], plan['spy']
        pd.a0, pd.a1, pd.fc2 = 2.0 * fin['f0'] / C, 2.0 * fin['df'] / C, 2.0 * fin['fc'] / C
        for k in range(3):
            pd.e1[k], pd.e2[k] = float(plan['e1'][k]), float(plan['e2'][k])
        pd.fu, pd.fr0, pd.fcen = p64(coll['final']['u']), p64(coll['final']['r0']), p64(fin['cen'])
        txy = np.ascontiguousarray(tile_digits(plan), np.int32)
        keep.append(txy)
        pd.tile_xy = txy.ctypes.data_as(_hip.P_i32)
        if window is not None:
            pd.wp, pd.wk = p32(window[0]), p32(window[1])
        if wf is not None:
            w0, gxv, gyr = self._weights(plan, wf)
            pd.w0, pd.gx, pd.gyr = p32(w0), p32(gxv), p32(gyr)
        h = ctypes.c_void_p()
        _hip.check(L.fsh_ffbp_create(ctypes.byref(pd), descs, ctypes.byref(h)), 'ffbp_create')
        self.h = h
        del keep

    @staticmethod
    def _weights(plan, wf):
        """Aperture weights [ntiles, Pf] (or [3, ntiles, Pf] with the e1/e2 gradient) in the final stage's form
        (w0, gx, gyr): (w0 + gx dx)(1 + gyr dy), gyr = gy / w0 where w0 is not negligible within the first-level
        tile. (fastsar/ffbp_cpu.py:222-229)"""
        wf = np.asarray(wf, np.float64)
        if wf.ndim == 2:
            return wf, np.zeros_like(wf), np.zeros_like(wf)
        G = plan['levels'][0]['C']
        w0 = wf[0].reshape(G, -1)
        big = np.abs(w0) > 1e-6 * np.maximum(np.abs(w0).max(1, keepdims=True), 1e-30)
        gyr = np.where(big, wf[2].reshape(G, -1) / np.where(w0 == 0, 1, w0), 0.0).reshape(wf[0].shape)
        return wf[0], wf[1], gyr

    def set_group(self, ng):
        """First-level children per group (0 or None: automatic, up to 8, halved while device memory lacks)."""
        _hip.check(_hip.lib().fsh_ffbp_set_group(self.h, int(ng or 0)), 'ffbp_set_group')
        self._group_req = int(ng or 0)

    def info(self):
        ng = ctypes.c_int()
        nb = ctypes.c_size_t()
        _hip.check(_hip.lib().fsh_ffbp_info(self.h, ctypes.byref(ng), ctypes.byref(nb)), 'ffbp_info')
        return dict(group=ng.value, device_bytes=nb.value)

    def adopt(self, donor):
        """Take over the group buffers of another HipFFBP whose plan has this one's shape (fsh_ffbp_adopt): -> True
        when taken (the donor must make its buffers again to form), False when the shapes differ or the donor has
        formed nothing yet."""
        got = ctypes.c_int(0)
        _hip.check(_hip.lib().fsh_ffbp_adopt(self.h, donor.h, ctypes.byref(got)), 'ffbp_adopt')
        return bool(got.value)

    def generation(self):
        """A counter that changes whenever the former makes its group buffers again."""
        g = ctypes.c_ulonglong()
        _hip.check(_hip.lib().fsh_ffbp_generation(self.h, ctypes.byref(g)), 'ffbp_generation')
        return g.value

    def level_forms(self):
        """How each level runs: 0 two passes, 1 one pass on VALU, 2 one pass on the matrix cores."""
        n = len(self.plan['levels'])
        v = (ctypes.c_int * n)()
        _hip.check(_hip.lib().fsh_ffbp_level_forms(self.h, v, n), 'ffbp_level_forms')
        return list(v[:])

    def set_profile(self, on):
        """Per-launch device timing (events on the former's stream, each launch waited for) of the next calls."""
        _hip.check(_hip.lib().fsh_ffbp_set_profile(self.h, int(bool(on))), 'ffbp_set_profile')

    def profile(self):
        """[(stage, level, ms)] per launch of the last profiled call; level == nlev for the final stage."""
        n = ctypes.c_int()
        _hip.check(_hip.lib().fsh_ffbp_profile(self.h, None, None, None, 0, ctypes.byref(n)), 'ffbp_profile')
        st, lv, ms = (ctypes.c_int * n.value)(), (ctypes.c_int * n.value)(), (ctypes.c_float * n.value)()
        _hip.check(_hip.lib().fsh_ffbp_profile(self.h, st, lv, ms, n.value, ctypes.byref(n)), 'ffbp_profile')
        return [(STAGES[st[i]], lv[i], float(ms[i])) for i in range(n.value)]

    def stage_ms(self):
        ms = (ctypes.c_float * len(STAGES))()
        _hip.check(_hip.lib().fsh_ffbp_stage_ms(self.h, ms, len(STAGES)), 'ffbp_stage_ms')
        return dict(zip(STAGES, [float(v) for v in ms]))

    def form16(self, h, out=None):
        """h: a TypedHistory (fastsar_amd.precision: 16-bit floats, CPHD integer samples), host or device, decoded
        on the device; out as in __call__."""
        if h.shape != (self.P, self.K):
            raise ValueError(f'phase history must be {(self.P, self.K)}, got {h.shape}')
        img, img_dev = _hip.output(out, (self.nx, self.ny))
        bad = ctypes.c_int(-1)
        data = h.data if h.on_device else np.ascontiguousarray(h.data)
        inv = None if h.inv is None else (h.inv if h.on_device else np.ascontiguousarray(h.inv, np.float32))
        _hip.check(_hip.lib().fsh_ffbp_form_typed(self.h, _hip.ptr(data), h.code, _hip.ptr(inv), int(h.on_device),
                                                  _hip.ptr(img), int(img_dev), ctypes.byref(bad)), 'ffbp_form')
        return img

    def __call__(self, S, out=None):
        """S [P, K] complex64 (host, C-contiguous) or a DeviceArray; out: None (a new host image), or a complex64
        [nx, ny] DeviceArray or numpy array to write into. Returns the image."""
        on_dev = isinstance(S, _hip.DeviceArray)
        if on_dev:
            if S.shape != (self.P, self.K) or S.dtype != np.complex64:
                raise ValueError(f'phase history must be complex64 {(self.P, self.K)}, got {S.dtype} {S.shape}')
        else:
            S = np.ascontiguousarray(S, np.complex64)
            if S.shape != (self.P, self.K):
                raise ValueError(f'phase history must be {(self.P, self.K)}, got {S.shape}')
        img, img_dev = _hip.output(out, (self.nx, self.ny))
        bad = ctypes.c_int(-1)
        _hip.check(_hip.lib().fsh_ffbp_form(self.h, _hip.ptr(S), int(on_dev), _hip.ptr(img), int(img_dev), ctypes.byref(bad)),
                   'ffbp_form')
        g = self.info()['group']
        if self._group_req == 0 and g < min(8, self.G) and g != self._warned_group:
            self._warned_group = g
            _mem.warn(f'hip: out of device memory for groups of {min(8, self.G)} first-level children; the former '
                      f'continues with {g} per group')
        return img

    def __del__(self):
        try:
            if getattr(self, 'h', None) and _hip._lib is not None:
                _hip._lib.fsh_ffbp_destroy(self.h)
        except Exception:
            pass
        self.h = None


def make_ffbp_hip(plan, coll, wf=None, window=None):
    """form(S) -> complex64 image [nx, ny] for the plan and collection arrays; window: (wp, wk) applied on the
    device, or None for a history already windowed."""
    return HipFFBP(plan, coll, wf=wf, window=window)
