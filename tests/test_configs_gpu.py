"""The full-size workloads on the GPU.

12,288 x 8,192 image from 16,384 pulses x 8,192 samples (6 degree aperture, 2,000 targets over a 6 km x 4 km scene
plus white clutter 25 dB down; T 8, four levels, 80 dB filters): error against the float64 device checker on seeded
256^2 tiles, the checker itself against the float64 direct sum over pulses and samples on the CPU at this geometry,
and the image of groups of 4 first-level children equal bit for bit to that of groups of 8 and to that formed with
complex values, not split operand words, handed from level 1 to level 2.

48 uneven images (sides 256 to 4,096 px, 512 to 8,192 pulses x 1,024 to 4,096 samples, spotlight and 2-degree
squint, seed 5): each image's final tile chosen by the predicted final-stage error (T 8, else 4, four levels; no
timing) with 4 dB kept in hand, since the prediction can read low by that much (image 11, 534^2 from 1400 x 4088:
-63.2 dB predicted at T 8, -59.2 dB measured, the same in float64 on the CPU and on the device), every replay of
the batch graph equal bit for bit to the images of single calls, and every image within -60 dB of the float64 checker
on tiles about two of its targets."""
import numpy as np
import pytest

import oracle
from fastsar_amd import ImageFormer, _hip, plan as pl, sim
from fastsar_amd.api import _window
from fastsar_amd.batch import Batch
from fastsar_amd.exact import bp_points_f64_device, simulate_device
from fastsar_amd.metrics import rel_db

pytestmark = pytest.mark.gpu


def _tile_points(grid, tiles, side):
    e1, e2 = np.asarray(grid['e1']), np.asarray(grid['e2'])
    pts = []
    for i0, j0 in tiles:
        I, J = np.meshgrid(np.arange(i0, i0 + side), np.arange(j0, j0 + side), indexing='ij')
        p = ((I - grid['nx'] / 2.0) * grid['spx'])[..., None] * e1 + ((J - grid['ny'] / 2.0) * grid['spy'])[..., None] * e2
        pts.append(p.reshape(-1, 3))
    return np.concatenate(pts)


def _checker(dS, col, pts, upsample=64):
    nfft = 1 << int(np.ceil(np.log2(upsample * col.K)))
    wp, wk = _window(col.Np, col.K)
    return bp_points_f64_device(dS, col.ant, np.linalg.norm(col.ant, axis=1), col.fmin, col.df, pts, nfft, True, wp, wk,
                                chunk=max(1, (1 << 30) // (nfft * 16)))


@pytest.fixture(scope='module')
def full():
    col = sim.spotlight(16384, 8192, 6.0)
    sp = 6000.0 / 12288
    grid = dict(nx=12288, ny=8192, spx=sp, spy=sp, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    rng = np.random.default_rng(2)
    pos, amp = sim.point_scene(rng, 2000, 6000.0, 4000.0, grid['e1'], grid['e2'])
    print('12288 x 8192 from 16384 x 8192: generating the history', flush=True)
    dS = simulate_device(col, pos, amp, clutter_sigma=sim.clutter_sigma(amp, -25.0), seed=2)
    return col, grid, pos, dS


@pytest.mark.timeout(600)
def test_full_image_error_and_groups(full, monkeypatch):
    """also: level 2 stages its rows from the bfloat16 x 3 words the level-1 pulse filter stores for it; with the
    complex values handed over instead (FSH_SPLIT_HANDOFF=0) the image is the same bit for bit"""
    col, grid, pos, dS = full
    f = ImageFormer(col.ant, col.fmin, col.df, col.K, **grid, T=8, levels=4, atten=80.0)
    img = _hip.DeviceArray((grid['nx'], grid['ny']), np.complex64)
    f(dS, out=img)
    assert f._form.info()['group'] == 8
    a = img.to_host()
    rng = np.random.default_rng(7)
    tiles = [(int(rng.integers(0, grid['nx'] - 255)), int(rng.integers(0, grid['ny'] - 255))) for _ in range(4)]
    e = rel_db(np.concatenate([a[i:i + 256, j:j + 256].ravel() for i, j in tiles]), _checker(dS, col, _tile_points(grid, tiles, 256)))
    print(f'  groups of 8: {e:.2f} dB against the float64 checker on 4 tiles of 256^2', flush=True)
    assert e <= -60.0
    f._form.set_group(4)
    f(dS, out=img)
    assert f._form.info()['group'] == 4
    assert np.array_equal(img.to_host(), a)
    monkeypatch.setenv('FSH_SPLIT_HANDOFF', '0')
    f(dS, out=img)
    assert np.array_equal(img.to_host(), a)


@pytest.mark.timeout(600)
def test_full_image_float16_rows(full):
    """precision='float16' at full size: level 1's range-filtered rows stored as scaled float16 pairs, 320 samples
    apart (rows of 315 padded to whole 128-byte lines), the image within 0.5 dB of the float32 one against the float64
    checker on the same tiles"""
    col, grid, pos, dS = full
    rng = np.random.default_rng(11)
    tiles = [(int(rng.integers(0, grid['nx'] - 255)), int(rng.integers(0, grid['ny'] - 255))) for _ in range(4)]
    ref = _checker(dS, col, _tile_points(grid, tiles, 256))
    img = _hip.DeviceArray((grid['nx'], grid['ny']), np.complex64)
    errs = {}
    for prec in ('float32', 'float16'):
        f = ImageFormer(col.ant, col.fmin, col.df, col.K, **grid, T=8, levels=4, atten=80.0, precision=prec)
        f(dS, out=img)
        a = img.to_host()
        errs[prec] = rel_db(np.concatenate([a[i:i + 256, j:j + 256].ravel() for i, j in tiles]), ref)
        del f
    print(f'  float32 {errs["float32"]:.2f} dB, float16 rows {errs["float16"]:.2f} dB on 4 tiles of 256^2', flush=True)
    assert errs['float32'] <= -60.0 and errs['float16'] <= errs['float32'] + 0.5


@pytest.mark.timeout(600)
def test_full_geometry_checker_against_direct_sum(full):
    """the float64 device checker (64x oversampled complex128 profiles, cubic reads) against the interpolation-free
    float64 sum over all 16384 x 8192 samples, on the CPU, at four target pixels and four seeded ones"""
    col, grid, pos, dS = full
    e1, e2 = np.asarray(grid['e1']), np.asarray(grid['e2'])
    near = [(int(round(t @ e1 / grid['spx'] + grid['nx'] / 2)), int(round(t @ e2 / grid['spy'] + grid['ny'] / 2)))
            for t in pos[:4]]
    rng = np.random.default_rng(8)
    near += [(int(rng.integers(0, grid['nx'])), int(rng.integers(0, grid['ny']))) for _ in range(4)]
    pts = np.array([((i - grid['nx'] / 2.0) * grid['spx']) * e1 + ((j - grid['ny'] / 2.0) * grid['spy']) * e2 for i, j in near])
    got = _checker(dS, col, pts)
    wp, wk = _window(col.Np, col.K)
    S = dS.to_host()
    want = np.zeros(len(pts), np.complex128)
    for p0 in range(0, col.Np, 1024):
        Sw = S[p0:p0 + 1024].astype(np.complex128) * (wp[p0:p0 + 1024, None] * wk[None, :]).astype(np.float64)
        want += oracle.direct_sum(Sw, col.ant[p0:p0 + 1024], col.fmin, col.df, pts, ref=np.linalg.norm(col.ant[p0:p0 + 1024], axis=1))
    e = oracle.plain_db(got, want)
    print(f'  checker against the direct sum at 8 pixels: {e:.1f} dB', flush=True)
    assert e < -90


@pytest.mark.timeout(900)
def test_uneven_batch_48_images():
    scenes = sim.uneven_batch(seed=5, n=48)
    formers, hists = [], []
    for i, sc in enumerate(scenes):
        col, g = sc['col'], sc['grid']
        fmax = col.fmin + col.K * col.df
        T, _ = pl.choose_T(col.ant, fmax, g['nx'], g['ny'], g['spx'], g['spy'], g['e1'], g['e2'], -64.0, choices=(8, 4))
        hists.append(simulate_device(col, sc['pos'], sc['amp'], clutter_sigma=sc['sigma'], seed=100 + i))
        formers.append(ImageFormer(col.ant, col.fmin, col.df, col.K, **g, T=T, levels=4, atten=80.0))
    print(f'48 images, sides {min(s["grid"]["nx"] for s in scenes)}-{max(s["grid"]["nx"] for s in scenes)}', flush=True)
    single = [f(S) for f, S in zip(formers, hists)]
    b = Batch(formers, hists, streams=4)
    for rep in range(2):
        r = b.run()
        assert r['bad'] == [-1] * 48
        for i, im in enumerate(b.images):
            assert np.array_equal(im.to_host(), single[i]), (rep, i)
        formers[rep](hists[rep])
        _hip.check(_hip.lib().fsh_synchronize())
    worst = -np.inf
    for i, (sc, S) in enumerate(zip(scenes, hists)):
        col, g = sc['col'], sc['grid']
        side = min(64, g['nx'])
        e1, e2 = np.asarray(g['e1']), np.asarray(g['e2'])
        tiles = []
        for t in sc['pos'][:2]:
            ci, cj = int(round(t @ e1 / g['spx'] + g['nx'] / 2)), int(round(t @ e2 / g['spy'] + g['ny'] / 2))
            tiles.append((min(max(ci - side // 2, 0), g['nx'] - side), min(max(cj - side // 2, 0), g['ny'] - side)))
        got = np.concatenate([single[i][a:a + side, c:c + side].ravel() for a, c in tiles])
        e = rel_db(got, _checker(S, col, _tile_points(g, tiles, side)))
        worst = max(worst, e)
        print(f'  image {i}: {g["nx"]}^2 from {col.Np} x {col.K}, T {formers[i].T}: {e:.1f} dB', flush=True)
        assert e <= -60.0, (i, e)
    print(f'  worst {worst:.1f} dB')
