"""HIP factorized backprojection against the float64 oracle: every stage on identical inputs, and whole images."""
import ctypes

import numpy as np
import pytest

import oracle
from fastsar_amd import _hip, plan as pl, sim
from fastsar_amd.ffbp_hip import make_ffbp_hip

pytestmark = pytest.mark.gpu


def db(a, b):
    return oracle.plain_db(a, b)


def cplx(re, im):
    return (np.asarray(re) + 1j * np.asarray(im)).astype(np.complex64)


ROT_CASES = [(3, 23, 5, 2, 301), (9, 67, 3, 1, 301), (4, 31, 2, 3, 301), (1, 1, 4, 1, 301), (9, 67, 64, 2, 2739),
             (3, 23, 17, 1, 4001), (8, 60, 24, 3, 313), (2, 18, 6, 2, 196), (120, 870, 2, 1, 12000)]


@pytest.mark.parametrize('kernel', ['valu', 'mfma', 'mfma1', 'long'])
@pytest.mark.parametrize('D,L,Cn,Np,K', ROT_CASES)
def test_rot_fir_k_stage(D, L, Cn, Np, K, kernel, monkeypatch):
    """rotation + range decimation on every kernel (VALU float32; the MFMA row kernel and the per-wave-row MFMA
    kernel, bfloat16 x 3; the direct form for windows longer than an LDS tile, which the last case, D = 120, takes
    by default) against the float64 oracle"""
    monkeypatch.setenv('FSH_ROT_KERNEL', kernel)
    rng = np.random.default_rng(D * 100 + L + Cn)
    P = 37
    x = (rng.standard_normal((Np, P, K)) + 1j * rng.standard_normal((Np, P, K))).astype(np.complex64)
    c0 = rng.uniform(-0.5, 0.5, (Np, Cn, P)).astype(np.float32)
    sl = rng.uniform(-0.8, 0.8, (Np, Cn, P)).astype(np.float32)
    if D == 1:
        fr, Ko, taps = None, K, np.ones(1, np.float32)
    else:
        fr, _ = pl.decimator_fir(K, D, 0.6)
        Ko, taps = fr['n_out'], fr['kern'].astype(np.float32)
    kc = (K - 1) / 2.0
    yre, yim = oracle.rot_fir_k(x.real, x.imag, c0, sl, fr, D, Ko, kc)
    dx, dc, ds, dt = (_hip.DeviceArray.from_host(a) for a in (x, c0, sl, taps))
    dy = _hip.DeviceArray((Np * Cn, P, Ko), np.complex64)
    _hip.check(_hip.lib().fsh_rot_fir_k(_hip.ptr(dx), Np, P, K, _hip.ptr(dc), _hip.ptr(ds), Cn, _hip.ptr(dt), len(taps),
                                        0 if fr is None else fr['pl'], D, Ko, kc, _hip.ptr(dy), None))
    got = dy.to_host()
    e = db(got, yre + 1j * yim)
    print(f'rot_fir_k {kernel} D={D} L={L} C={Cn} K={K}: {e:.1f} dB')
    assert e < (-120 if kernel in ('valu', 'long') else -100), e


@pytest.mark.parametrize('kernel', ['stream', 'lds'])
@pytest.mark.parametrize('D,pf', [(2, 0.5), (11, 0.7), (5, 0.3), (3, 0.9), (8, 0.6)])
def test_fir_p_stage(D, pf, kernel, monkeypatch):
    """pulse decimation (streaming sliding-window kernel and the LDS-staged one) against the float64 oracle"""
    monkeypatch.setenv('FSH_FIRP_KERNEL', kernel)
    rng = np.random.default_rng(D)
    B, P, Ko = 3, 613, 70
    fr, _ = pl.decimator_fir(P, D, pf)
    y = (rng.standard_normal((B, P, Ko)) + 1j * rng.standard_normal((B, P, Ko))).astype(np.complex64)
    zre, zim = oracle.fir_p(y.real, y.imag, fr, D, fr['n_out'])
    dy, dt = _hip.DeviceArray.from_host(y), _hip.DeviceArray.from_host(fr['kern'].astype(np.float32))
    dz = _hip.DeviceArray((B, fr['n_out'], Ko), np.complex64)
    _hip.check(_hip.lib().fsh_fir_p(_hip.ptr(dy), B, P, Ko, _hip.ptr(dt), fr['L'], fr['pl'], D, fr['n_out'], _hip.ptr(dz), None))
    e = db(dz.to_host(), zre + 1j * zim)
    print(f'fir_p D={D} L={fr["L"]}: {e:.1f} dB')
    assert e < -125, e


FUSED_CASES = [(2, 0.6, 3, 0.6, 6, 2, 198, 50), (4, 0.6, 4, 0.6, 5, 1, 300, 315), (1, 0, 2, 0.5, 3, 2, 100, 40),
               (3, 0.5, 1, 0, 4, 1, 77, 90), (4, 0.6, 3, 0.7, 17, 3, 150, 200), (2, 0.4, 4, 0.7, 8, 1, 613, 30),
               (2, 0.5, 2, 0.5, 4, 9, 143, 183)]


@pytest.mark.parametrize('Dk,pk,Dp,pp,Cn,Np,P,K', FUSED_CASES)
def test_level_fused_stage(Dk, pk, Dp, pp, Cn, Np, P, K):
    """a level's rotation and both filters in one pass against the float64 oracle's two stages"""
    rng = np.random.default_rng(Dk * 1000 + Dp * 100 + Cn)
    x = (rng.standard_normal((Np, P, K)) + 1j * rng.standard_normal((Np, P, K))).astype(np.complex64)
    c0 = rng.uniform(-0.5, 0.5, (Np, Cn, P)).astype(np.float32)
    sl = rng.uniform(-0.8, 0.8, (Np, Cn, P)).astype(np.float32)
    frk = pl.decimator_fir(K, Dk, pk)[0] if Dk > 1 else None
    frp = pl.decimator_fir(P, Dp, pp)[0] if Dp > 1 else None
    Ko = frk['n_out'] if frk else K
    Po = frp['n_out'] if frp else P
    kc = (K - 1) / 2.0
    yre, yim = oracle.rot_fir_k(x.real, x.imag, c0, sl, frk, Dk, Ko, kc)
    if frp:
        zre, zim = oracle.fir_p(yre, yim, frp, Dp, Po)
    else:
        zre, zim = yre, yim
    tk = frk['kern'].astype(np.float32) if frk else np.ones(1, np.float32)
    tp = frp['kern'].astype(np.float32) if frp else np.ones(1, np.float32)
    dx, dc, ds, dtk, dtp = (_hip.DeviceArray.from_host(a) for a in (x, c0, sl, tk, tp))
    dz = _hip.DeviceArray((Np * Cn, Po, Ko), np.complex64)
    print(f'level_fused Dk={Dk} Dp={Dp} C={Cn} Np={Np} P={P} K={K} -> [{Po} x {Ko}]', flush=True)
    _hip.check(_hip.lib().fsh_level_fused(_hip.ptr(dx), Np, P, K, _hip.ptr(dc), _hip.ptr(ds), Cn, _hip.ptr(dtk), len(tk),
                                          frk['pl'] if frk else 0, Dk, Ko, _hip.ptr(dtp), len(tp), frp['pl'] if frp else 0,
                                          Dp, Po, kc, _hip.ptr(dz), None), 'level_fused')
    e = db(dz.to_host(), zre + 1j * zim)
    print(f'  {e:.1f} dB')
    assert e < -115, e


MFMA_LEVEL_CASES = [(9, 0.5, 11, 0.6, 64, 1, 300, 600), (8, 0.6, 8, 0.6, 64, 2, 200, 315), (3, 0.5, 4, 0.7, 13, 2, 150, 200),
                    (4, 0.6, 4, 0.6, 8, 1, 2000, 100), (2, 0.6, 2, 0.6, 8, 1, 100, 700), (3, 0.4, 6, 0.6, 24, 3, 260, 90)]


@pytest.mark.parametrize('Dk,pk,Dp,pp,Cn,Np,P,K', MFMA_LEVEL_CASES)
def test_level_mfma_stage(Dk, pk, Dp, pp, Cn, Np, P, K):
    """a level on the matrix cores with the pulse filter fused (bfloat16 x 3 range filter, float32 pulse filter)
    against the float64 oracle's two stages"""
    rng = np.random.default_rng(Dk * 1000 + Dp * 100 + Cn + K)
    x = (rng.standard_normal((Np, P, K)) + 1j * rng.standard_normal((Np, P, K))).astype(np.complex64)
    c0 = rng.uniform(-0.5, 0.5, (Np, Cn, P)).astype(np.float32)
    sl = rng.uniform(-0.8, 0.8, (Np, Cn, P)).astype(np.float32)
    frk, _ = pl.decimator_fir(K, Dk, pk)
    frp, _ = pl.decimator_fir(P, Dp, pp)
    Ko, Po, kc = frk['n_out'], frp['n_out'], (K - 1) / 2.0
    yre, yim = oracle.rot_fir_k(x.real, x.imag, c0, sl, frk, Dk, Ko, kc)
    zre, zim = oracle.fir_p(yre, yim, frp, Dp, Po)
    dx, dc, ds = (_hip.DeviceArray.from_host(a) for a in (x, c0, sl))
    dtk, dtp = (_hip.DeviceArray.from_host(f['kern'].astype(np.float32)) for f in (frk, frp))
    dz = _hip.DeviceArray((Np * Cn, Po, Ko), np.complex64)
    print(f'level_mfma Dk={Dk} Lk={frk["L"]} Dp={Dp} Lp={frp["L"]} C={Cn} Np={Np} P={P} K={K} -> [{Po} x {Ko}]', flush=True)
    _hip.check(_hip.lib().fsh_level_mfma(_hip.ptr(dx), Np, P, K, _hip.ptr(dc), _hip.ptr(ds), Cn, _hip.ptr(dtk), frk['L'],
                                         frk['pl'], Dk, Ko, _hip.ptr(dtp), frp['L'], frp['pl'], Dp, Po, kc, _hip.ptr(dz), None),
               'level_mfma')
    e = db(dz.to_host(), zre + 1j * zim)
    print(f'  {e:.1f} dB')
    # bfloat16 x 3 leaves out the lo x lo product, 2^-16 of each term: about -96 dB on long filters
    assert e < -94, e


def test_level_fused_rejects():
    """shapes outside the fused form are refused before any launch"""
    lib = _hip.lib()
    dz = _hip.DeviceArray((1,), np.complex64)
    p = _hip.ptr(dz)
    assert lib.fsh_level_fused(p, 1, 10, 20, p, p, 1, p, 25, 12, 5, 8, p, 3, 1, 2, 6, 9.5, p, None) == 2   # Dk = 5
    assert lib.fsh_level_fused(p, 1, 10, 600, p, p, 1, None, 1, 0, 1, 600, None, 1, 0, 1, 10, 9.5, p, None) == 2   # Ko > 256


def _final_inputs(rng, B, Pf, Qf, T):
    d = (rng.standard_normal((B, Pf, Qf)) + 1j * rng.standard_normal((B, Pf, Qf))).astype(np.complex64)
    ux = rng.uniform(-0.1, 0.1, (B, Pf)) + 0.05
    uy = rng.uniform(0.8, 0.9, (B, Pf))
    geo = np.stack([rng.uniform(0, .1, B), rng.uniform(.8, .9, B), rng.uniform(9e3, 1.1e4, B)], 1)
    return d, ux, uy, geo


@pytest.mark.parametrize('weights', [False, True])
@pytest.mark.parametrize('kernel', ['reg', 'lds'])
@pytest.mark.parametrize('T', [4, 8, 16, 32])
def test_final_tiles_stage(T, kernel, weights, monkeypatch):
    """final stage (register-blocked and LDS-table kernels, with and without aperture weights) against float64"""
    monkeypatch.setenv('FSH_FINAL_KERNEL', kernel)
    rng = np.random.default_rng(T)
    B, Pf, Qf = 5, 23, 41
    spx, spy = 0.5, 0.45
    f0, df = 9.45e9, 3e6
    a0, a1, fc2 = 2 * f0 / pl.C, 2 * df / pl.C, 2 * (f0 + (Qf - 1) / 2 * df) / pl.C
    d, ux, uy, geo = _final_inputs(rng, B, Pf, Qf, T)
    dlx = (np.arange(T) - (T - 1) / 2.0) * spx
    dly = (np.arange(T) - (T - 1) / 2.0) * spy
    wts = None
    if weights:
        wts = (rng.uniform(0.5, 1.5, (B, Pf)), rng.uniform(-0.01, 0.01, (B, Pf)), rng.uniform(-0.01, 0.01, (B, Pf)))
        wts = tuple(w.astype(np.float32) for w in wts)
    ore, oim = oracle.final_tiles(d.real, d.imag, ux, uy, dlx, dly, a0, a1, geo[:, 0], geo[:, 1], geo[:, 2], fc2, T, wts)
    txy = np.stack([np.arange(B) * T, np.zeros(B)], 1).astype(np.int32)
    dd, dux, duy, dgeo, dtxy = (_hip.DeviceArray.from_host(a) for a in (d, ux, uy, geo, txy))
    dw = [None] * 3 if wts is None else [_hip.DeviceArray.from_host(w) for w in wts]
    out = _hip.DeviceArray((B * T, T), np.complex64)
    _hip.check(_hip.lib().fsh_final_tiles(_hip.ptr(dd), B, Pf, Qf, T, _hip.ptr(dux), _hip.ptr(duy), _hip.ptr(dgeo), spx, spy,
                                          a0, a1, fc2, *[_hip.ptr(w) for w in dw], _hip.ptr(dtxy), 0, 0, B * T, T,
                                          _hip.ptr(out), None))
    e = db(out.to_host().reshape(B, T, T), ore + 1j * oim)
    print(f'final_tiles {kernel} T={T} weights={weights}: {e:.1f} dB')
    assert e < -110, e


def test_device_phases_stage():
    rng = np.random.default_rng(4)
    Np, Cn, P = 3, 6, 257
    ant = np.stack([np.full(P, -8660.0), np.linspace(-200, 200, P), np.full(P, 5000.0)], 1)
    r0 = np.linalg.norm(ant, axis=1)
    u = ant / r0[:, None]
    refs = rng.uniform(-300, 300, (Np, 3)) * [1, 1, 0]
    d = rng.uniform(-40, 40, (Cn, 3)) * [1, 1, 0]
    k0c, k1 = 64.0, 2.4e-4
    c0, sl = pl.device_phases(u, r0, d, refs, k0c, k1)
    args = [_hip.DeviceArray.from_host(a) for a in (refs, u, r0, d)]
    dc, ds = _hip.DeviceArray((Np, Cn, P), np.float32), _hip.DeviceArray((Np, Cn, P), np.float32)
    _hip.check(_hip.lib().fsh_device_phases(*[_hip.ptr(a) for a in args], Np, Cn, P, k0c, k1, _hip.ptr(dc), _hip.ptr(ds), None))
    gc, gs = dc.to_host(), ds.to_host()
    dphi = np.abs((gc - c0 + 0.5) % 1.0 - 0.5)
    print(f'device_phases: max phase diff {dphi.max():.2e} cycles, slope {np.abs(gs - sl).max():.2e}')
    assert dphi.max() < 1e-6 and np.allclose(gs, sl, rtol=1e-6, atol=0)


@pytest.fixture(scope='module')
def scene():
    rng = np.random.default_rng(1)
    col = sim.make_collect(res=0.5, scene=60.0, r0=5e3)
    pos = np.stack([rng.uniform(-25, 25, 40), rng.uniform(-25, 25, 40), np.zeros(40)], 1)
    amp = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    return col, sim.simulate_brute(col, pos, amp).astype(np.complex64)


@pytest.mark.parametrize('kernel', ['valu', 'mfma', 'fused', 'auto', 'lmfma', 'rot-in-range', 'rows1', 'rows8', 'cols1',
                                    'cols2'])
@pytest.mark.parametrize('nlev,T', [(2, 16), (3, 16), (3, 8), (2, 32), (4, 4)])
def test_image_vs_oracle(scene, nlev, T, kernel, monkeypatch):
    """whole images against the float64 oracle's image: float32 VALU levels, bfloat16 x 3 MFMA levels, MFMA level 0
    with the later levels fused on VALU where they fit, the default choice, and the later levels fused on the matrix
    cores where they fit; on the two-pass MFMA levels also with the range rotation kept in the range filter instead
    of the pulse filter, and with the pulse filter on one column per thread (its rows loaded one, four or eight at a
    time) or two"""
    if kernel == 'lmfma':
        monkeypatch.setenv('FSH_LEVEL_KERNEL', 'mfma')
    elif kernel in ('rot-in-range', 'rows1', 'rows8', 'cols1', 'cols2'):
        monkeypatch.setenv('FSH_ROT_KERNEL', 'mfma')
        monkeypatch.setenv('FSH_LEVEL_KERNEL', 'split')
        if kernel == 'rot-in-range':
            monkeypatch.setenv('FSH_DEFER_ROT', '0')
        elif kernel in ('cols1', 'cols2'):
            monkeypatch.setenv('FSH_FIRP_COLS', kernel[4:])
        else:
            monkeypatch.setenv('FSH_FIRP_COLS', '1')
            monkeypatch.setenv('FSH_FIRP_ROWS', kernel[4:])
    elif kernel != 'auto':
        monkeypatch.setenv('FSH_ROT_KERNEL', 'valu' if kernel == 'valu' else 'mfma')
        monkeypatch.setenv('FSH_LEVEL_KERNEL', 'fused' if kernel == 'fused' else 'split')
    col, S = scene
    e1, e2 = (0.0, 1.0, 0.0), (1.0, 0.0, 0.0)
    p = pl.make_plan(col, 128, 128, 0.5, 0.5, T=T, nlev=nlev, e1=e1, e2=e2)
    coll = pl.collection_arrays(p, col.ant)
    ref = oracle.ffbp_image(p, coll, S, 'oracle')
    form = make_ffbp_hip(p, coll)
    img = form(S)
    e = db(img, ref)
    print(f'image {kernel} nlev={nlev} T={T}: hip vs float64 oracle {e:.1f} dB')
    assert e < (-105 if kernel == 'valu' else -95), e
    # groups of first-level children: the same bits whatever the group
    form.set_group(1)
    img1 = form(S)
    assert np.array_equal(img, img1)


def _group_bytes(plan, ng):
    """the largest level buffer of a group of ng first-level children on the two-pass path (bytes)"""
    big, parents = 0, 1
    for l, lv in enumerate(plan['levels']):
        outs = parents * (ng if l == 0 else lv['C'])
        big = max(big, outs * lv['P'] * lv['Ko'] * 8, outs * lv['Po'] * lv['Ko'] * 8)
        parents = outs
    return big


def test_group_halving_on_allocation_failure(scene, monkeypatch):
    """a device that cannot hold the buffers of 8 first-level children: the former halves the group until they
    fit, warns once (MemoryWarning), and forms the same image bit for bit"""
    from fastsar_amd.memory import MemoryWarning
    monkeypatch.setenv('FSH_LEVEL_KERNEL', 'split')
    col, S = scene
    p = pl.make_plan(col, 128, 128, 0.5, 0.5, T=8, nlev=2, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    coll = pl.collection_arrays(p, col.ant)
    assert p['levels'][0]['C'] >= 8
    want = make_ffbp_hip(p, coll)(S)
    form = make_ffbp_hip(p, coll)
    form.set_group(1)
    form(S)                                        # the history's upload buffer exists from here on
    form.set_group(0)
    cap = _group_bytes(p, 2) + 1                   # groups of 2 fit, groups of 4 do not
    assert _group_bytes(p, 4) > cap
    monkeypatch.setenv('FSH_FAIL_ALLOC_ABOVE', str(cap))
    with pytest.warns(MemoryWarning, match='continues with 2'):
        img = form(S)
    assert form.info()['group'] == 2
    assert np.array_equal(img, want)
    # a requested group that does not fit is an error, not a fallback
    form.set_group(8)
    with pytest.raises(MemoryError):
        form(S)
