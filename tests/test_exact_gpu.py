"""HIP exact backprojection (ExactFormer, backproject), the float64 device checker and the device scene generator
against the float64 oracle. Thresholds are the reference's (tests/test_exact.py:40-50): cubic < -65 dB and
linear < -52 dB against exact backprojection."""
import numpy as np
import pytest

import oracle
from fastsar_amd import _hip, sim
from fastsar_amd.exact import ExactFormer, backproject, bp_points_f64_device, simulate_device

pytestmark = pytest.mark.gpu

GRID = dict(nx=100, ny=120, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))


def _scene(r0, n=30, seed=5):
    rng = np.random.default_rng(seed)
    col = sim.make_collect(res=0.5, scene=60.0, r0=r0)
    tg = np.stack([rng.uniform(-25, 25, n), rng.uniform(-25, 25, n), np.zeros(n)], 1)
    amp = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return col, tg, amp, sim.simulate_brute(col, tg, amp).astype(np.complex64)


@pytest.fixture(scope='module', params=[20e3, 600e3], ids=['20km', '600km'])
def scene(request):
    col, tg, amp, S = _scene(request.param)
    wp, wk = oracle.window(*S.shape)
    ref = oracle.bp_points(S * wp[:, None] * wk[None, :], col.ant, col.fmin, col.df, oracle.plane_points(**GRID),
                           upsample=64, interp='cubic')
    return request.param, col, S, ref.reshape(GRID['nx'], GRID['ny'])


@pytest.mark.parametrize('interp,limit', [('cubic', -65.0), ('linear', -52.0)])
def test_exact_former_vs_oracle(scene, interp, limit):
    r0, col, S, ref = scene
    print(f'exact former {interp} at {r0 / 1e3:.0f} km', flush=True)
    f = ExactFormer(col.ant, col.fmin, col.df, S.shape[1], **GRID, interp=interp)
    img = f(S)
    e = oracle.rel_db(img, ref)
    print(f'  tile {f.tile}: {e:.1f} dB')
    assert img.shape == (GRID['nx'], GRID['ny']) and img.dtype == np.complex64
    assert e < limit, e


def test_exact_former_vs_float64_tiles(scene):
    """the same tiled algorithm in float64 on the same host plan: the device's float32 per-pixel arithmetic and the
    hipFFT profiles against the float64 restatement"""
    r0, col, S, _ = scene
    f = ExactFormer(col.ant, col.fmin, col.df, S.shape[1], **GRID)
    wp, wk = oracle.window(*S.shape)
    rc = oracle.range_compress(S * wp[:, None] * wk[None, :], f.nfft)
    win = np.take_along_axis(rc, (f.blo[:, None] + np.arange(f.W)[None, :]) % f.nfft, axis=1)
    want = oracle.exact_tiles(win, f.lo, col.ant, np.linalg.norm(col.ant, axis=1), f.cen, f.ntx, f.nty, f.tx, f.ty,
                              GRID['nx'], GRID['ny'], 0.5, 0.5, GRID['e1'], GRID['e2'], f.inv_dr, f.kcyc, True)
    e = oracle.plain_db(f(S), want)
    print(f'exact former vs float64 tiles at {r0 / 1e3:.0f} km: {e:.1f} dB')
    # float32 per-pixel phase (about 2^-24 of the few hundred cycles a tile spans) bounds the agreement
    assert e < -70, e


@pytest.mark.parametrize('r0,upsample,interp', [(250.0, 8, 'cubic'), (20e3, 64, 'cubic'), (20e3, 64, 'linear')],
                         ids=['250m-4x4-tiles', '20km-upsample64-16px-regions', '20km-upsample64-linear'])
def test_exact_region_shapes(r0, upsample, interp):
    """the region kernel's other shapes against the float64 tiled restatement on the same host plan: 4 x 4 expansion
    tiles near the antenna (16 tiles per 32-pixel region), and windows long enough (upsample 64) that the regions
    shrink to 16 x 16 pixels"""
    col, _, _, S = _scene(r0)
    f = ExactFormer(col.ant, col.fmin, col.df, S.shape[1], **GRID, upsample=upsample, interp=interp)
    print(f'exact former at {r0:.0f} m, upsample {upsample}, {interp}: tile {f.tile}, W {f.W}', flush=True)
    if r0 < 1e3:
        assert f.tile == (4, 4)
    wp, wk = oracle.window(*S.shape)
    rc = oracle.range_compress(S * wp[:, None] * wk[None, :], f.nfft)
    win = np.take_along_axis(rc, (f.blo[:, None] + np.arange(f.W)[None, :]) % f.nfft, axis=1)
    want = oracle.exact_tiles(win, f.lo, col.ant, np.linalg.norm(col.ant, axis=1), f.cen, f.ntx, f.nty, f.tx, f.ty,
                              GRID['nx'], GRID['ny'], 0.5, 0.5, GRID['e1'], GRID['e2'], f.inv_dr, f.kcyc, interp == 'cubic')
    e = oracle.plain_db(f(S), want)
    print(f'  {e:.1f} dB against the float64 tiles')
    assert e < -70, e


def test_exact_former_chunks_and_device_io(scene):
    """the image does not depend on the pulse chunk, and device input/output give the host result"""
    _, col, S, _ = scene
    a = ExactFormer(col.ant, col.fmin, col.df, S.shape[1], **GRID, chunk=1024)(S)
    b = ExactFormer(col.ant, col.fmin, col.df, S.shape[1], **GRID, chunk=37)(S)
    assert oracle.plain_db(b, a) < -120
    f = ExactFormer(col.ant, col.fmin, col.df, S.shape[1], **GRID)
    out = _hip.DeviceArray((GRID['nx'], GRID['ny']), np.complex64)
    f(_hip.DeviceArray.from_host(S), out=out)
    assert np.array_equal(out.to_host(), f(S))


@pytest.mark.parametrize('profiles,interp', [('float16', 'linear'), ('float16-pair', 'linear'), ('float16', 'cubic')])
def test_exact_float16_profiles(scene, profiles, interp):
    """float16 profiles (each chunk scaled to peak 2048) against float32 ones: the rounding (about -78 dB) stays
    below -60 dB (the reference's bound, tests/test_pfa_paths.py:280-291); with linear reads it is far below the
    interpolation error (the error against the FP64 oracle moves by less than 1 dB), with cubic reads it becomes the
    larger term (-76 dB against -81 dB in float32) and the image keeps the cubic bound, -65 dB
    (tests/test_exact.py:40-50). Chunks of 37 pulses, so that each chunk has its own scale."""
    r0, col, S, ref = scene
    kw = dict(interp=interp, chunk=37)
    a = ExactFormer(col.ant, col.fmin, col.df, S.shape[1], **GRID, profiles='float32', **kw)(S)
    f = ExactFormer(col.ant, col.fmin, col.df, S.shape[1], **GRID, profiles=profiles, **kw)
    assert f.profiles == profiles
    b = f(S)
    e, ea, eb = oracle.rel_db(b, a), oracle.rel_db(a, ref), oracle.rel_db(b, ref)
    print(f'{profiles} {interp} at {r0 / 1e3:.0f} km: {e:.1f} dB against float32 profiles; oracle {eb:.1f} (float32 '
          f'{ea:.1f}) dB')
    assert e < -60 and (eb < ea + 1.0 if interp == 'linear' else eb < -65), (e, ea, eb)
    assert abs(np.vdot(b, a) / np.vdot(a, a) - 1) < 1e-3          # the scale is undone exactly


def test_exact_profiles_choice():
    """'auto' is float32 on this backend, also for linear reads on grids of 2^25 pixels or more, where the reference
    picks float16 for its cuda kernel (fastsar/exact.py:420-430: float16 profiles are slower on gfx950); float16
    pairs are for linear reads only"""
    wide = sim.spotlight(64, 16384, 3.0, r0=20e3)                   # 8.2 km unambiguous range: a 4 km x 2 km grid fits
    big = dict(GRID, nx=8192, ny=4096)
    assert ExactFormer(wide.ant, wide.fmin, wide.df, wide.K, **big, interp='linear').profiles == 'float32'
    assert ExactFormer(wide.ant, wide.fmin, wide.df, wide.K, **big, interp='linear', profiles='float16').profiles == 'float16'
    col = sim.make_collect(res=0.5, scene=60.0, r0=20e3)
    assert ExactFormer(col.ant, col.fmin, col.df, col.K, **GRID, interp='linear').profiles == 'float32'
    with pytest.raises(ValueError, match='profiles'):
        ExactFormer(col.ant, col.fmin, col.df, col.K, **GRID, profiles='float8')
    with pytest.raises(ValueError, match='float16-pair'):
        ExactFormer(col.ant, col.fmin, col.df, col.K, **GRID, profiles='float16-pair')


def test_exact_former_rejects_nonfinite():
    col, _, _, S = _scene(20e3, n=3)
    S = S.copy()
    S[17, 5] = np.nan
    f = ExactFormer(col.ant, col.fmin, col.df, S.shape[1], **GRID)
    with pytest.raises(ValueError, match='17'):
        f(S)


@pytest.mark.parametrize('interp', ['linear', 'cubic'])
def test_bp_points_f64_device(interp):
    """the device float64 checker against the CPU float64 oracle on the same profile length: unwindowed, agreement
    to float64 rounding (the FFT libraries differ); windowed, to the float32 rounding of the window product, which
    the two sides form in different orders"""
    col, tg, amp, S = _scene(5e3, n=12, seed=2)
    rng = np.random.default_rng(9)
    pts = np.stack([rng.uniform(-30, 30, 500), rng.uniform(-30, 30, 500), rng.uniform(-1, 1, 500)], 1)
    nfft = 1 << int(np.ceil(np.log2(16 * S.shape[1])))
    ref = np.linalg.norm(col.ant, axis=1)
    dS = _hip.DeviceArray.from_host(S)
    got = bp_points_f64_device(dS, col.ant, ref, col.fmin, col.df, pts, nfft, interp == 'cubic', chunk=100)
    want = oracle.bp_points(S, col.ant, col.fmin, col.df, pts, upsample=16, interp=interp)
    e = oracle.plain_db(got, want)
    print(f'bp_points_f64 {interp}: {e:.1f} dB')
    assert e < -180, e
    b = backproject(S, col.ant, col.fmin, col.df, pts, upsample=16, interp=interp, chunk=64, window=False)
    assert oracle.plain_db(b, want) < -180
    wp, wk = oracle.window(*S.shape)
    got = bp_points_f64_device(dS, col.ant, ref, col.fmin, col.df, pts, nfft, interp == 'cubic', wp, wk, chunk=100)
    want = oracle.bp_points(S * wp[:, None] * wk[None, :], col.ant, col.fmin, col.df, pts, upsample=16, interp=interp)
    e = oracle.plain_db(got, want)
    print(f'  windowed: {e:.1f} dB')
    assert e < -140, e


def test_sim_points_vs_brute():
    """the device scene generator against the float64 direct sum (stored complex64)"""
    col, tg, amp, _ = _scene(5e3, n=20, seed=4)
    want = sim.simulate_brute(col, tg, amp)
    got = simulate_device(col, tg, amp).to_host()
    e = oracle.plain_db(got, want)
    print(f'sim_points: {e:.1f} dB')
    assert e < -130, e
    # clutter: white, of the requested power, reproducible per seed
    a = simulate_device(col, tg, amp, clutter_sigma=0.5, seed=11).to_host() - got
    b = simulate_device(col, tg, amp, clutter_sigma=0.5, seed=11).to_host() - got
    assert np.array_equal(a, b)
    assert abs(np.mean(np.abs(a) ** 2) / 0.25 - 1) < 0.02
