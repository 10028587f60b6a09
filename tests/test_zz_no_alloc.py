"""The no-allocation level entry point (fsh_level_mfma_noalloc) against fsh_level_mfma: the same images bit for bit on
two calls in a row, the second on a smaller level (fewer samples and a smaller polyphase table in the shared fixed
table), no device memory taken by the library, and a level whose table would not fit refused before any launch."""
import ctypes

import numpy as np
import pytest

from fastsar_amd import _hip, plan as pl

pytestmark = pytest.mark.gpu

# (Dk, range pass band, Dp, pulse pass band, children, parents, P, K): the second smaller than the first
CASES = [(9, 0.5, 11, 0.6, 64, 1, 300, 600), (8, 0.6, 8, 0.6, 16, 2, 200, 315)]


def _inputs(Dk, pk, Dp, pp, Cn, Np, P, K):
    rng = np.random.default_rng(Dk * 1000 + Dp * 100 + Cn + K)
    x = (rng.standard_normal((Np, P, K)) + 1j * rng.standard_normal((Np, P, K))).astype(np.complex64)
    c0 = rng.uniform(-0.5, 0.5, (Np, Cn, P)).astype(np.float32)
    sl = rng.uniform(-0.8, 0.8, (Np, Cn, P)).astype(np.float32)
    frk, _ = pl.decimator_fir(K, Dk, pk)
    frp, _ = pl.decimator_fir(P, Dp, pp)
    return x, c0, sl, frk, frp


def _device(x, c0, sl, frk, frp, Cn):
    Np, P, K = x.shape
    dev = [_hip.DeviceArray.from_host(a) for a in (x, c0, sl)]
    dev += [_hip.DeviceArray.from_host(f['kern'].astype(np.float32)) for f in (frk, frp)]
    return dev


def _run(name, dev, frk, frp, Dk, Dp, Cn, dz):
    dx, dc, ds, dtk, dtp = dev
    Np, P, K = dx.shape
    Ko, Po, kc = frk['n_out'], frp['n_out'], (K - 1) / 2.0
    _hip.check(getattr(_hip.lib(), name)(_hip.ptr(dx), Np, P, K, _hip.ptr(dc), _hip.ptr(ds), Cn, _hip.ptr(dtk), frk['L'],
                                         frk['pl'], Dk, Ko, _hip.ptr(dtp), frp['L'], frp['pl'], Dp, Po, kc, _hip.ptr(dz),
                                         None), name)


def _held():
    n = ctypes.c_size_t()
    _hip.check(_hip.lib().fsh_device_bytes(ctypes.byref(n)))
    return n.value


def test_noalloc_matches_level_mfma_twice():
    lib = _hip.lib()
    for Dk, pk, Dp, pp, Cn, Np, P, K in CASES:
        x, c0, sl, frk, frp = _inputs(Dk, pk, Dp, pp, Cn, Np, P, K)
        print(f'level Dk={Dk} Lk={frk["L"]} Dp={Dp} Lp={frp["L"]} C={Cn} Np={Np} P={P} K={K}', flush=True)
        dev = _device(x, c0, sl, frk, frp, Cn)
        want, got = (_hip.DeviceArray((Np * Cn, frp['n_out'], frk['n_out']), np.complex64) for _ in range(2))
        _run('fsh_level_mfma', dev, frk, frp, Dk, Dp, Cn, want)
        held = _held()
        _run('fsh_level_mfma_noalloc', dev, frk, frp, Dk, Dp, Cn, got)
        assert _held() == held                 # no device memory taken by the library
        assert np.array_equal(got.to_host(), want.to_host())
    _hip.check(lib.fsh_synchronize())


def test_noalloc_refuses_a_table_beyond_its_size():
    """Dp 300 with 16 pulse taps per phase needs a 4800-float table: refused with FSH_EINVAL, nothing launched"""
    lib = _hip.lib()
    d = _hip.DeviceArray((1,), np.complex64)
    p = _hip.ptr(d)
    Dk, Lk, K = 2, 8, 64
    Dp, Lp, P = 300, 300 * 15 + 1, 6000
    Ko, Po = (K + Dk - 1) // Dk + 8, (P + Dp - 1) // Dp + 32
    st = lib.fsh_level_mfma_noalloc(p, 1, P, K, p, p, 1, p, Lk, 3, Dk, Ko, p, Lp, 2000, Dp, Po, (K - 1) / 2.0, p, None)
    assert st == 2
    assert 'exceeds the fixed table' in _hip.last_error()
