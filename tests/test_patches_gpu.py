"""Patch mosaics on the GPU (fastsar_amd.patches.form_mosaic; fastsar/patches.py:619-970).

Dropped pulses: the reference's fixture (tests/test_patches.py:236-261 there), a spotlight at 5 km over a 600 m
scene with 80 point targets (seed 2), 3 gaps of 5 to 10 pulses and 3 runs of every other pulse dropped; the mosaic of
128 x 128 pixels at 0.5 m in patches of 64 x 64 against float64 exact backprojection of the remaining pulses
(-45 dB, the reference's bar; here against cubic reads of 64x oversampled profiles). The same mosaic with each patch
forming its own buffers is equal bit for bit to the one whose patches reuse them. A history whose pulses are
referenced to ranges other than |a_p| (ref + a smooth per-pulse offset, the history re-referenced accordingly) forms
the same mosaic (-60 dB, the reference's bar for its moving reference point). An azimuth window applied through a
beam against float64 backprojection with the same per-pixel weights."""
import numpy as np
import pytest

import oracle
from fastsar_amd import patches as pt, sim
from fastsar_amd.exact import bp_points_f64_device, simulate_device
from fastsar_amd.metrics import rel_db
from fastsar_amd.plan import C

pytestmark = pytest.mark.gpu

E1, E2 = np.array([0.0, 1.0, 0.0]), np.array([1.0, 0.0, 0.0])
ORIGIN = -32.0 * E1 - 32.0 * E2


def _grid_points(n=128, sp=0.5):
    i, j = np.meshgrid(np.arange(n) * sp, np.arange(n) * sp, indexing='ij')
    return ORIGIN + i[..., None] * E1 + j[..., None] * E2


@pytest.fixture(scope='module')
def dropped():
    rng = np.random.default_rng(2)
    col = sim.make_collect(res=0.5, scene=600.0, r0=5e3)
    tg = np.stack([rng.uniform(-25, 25, 80), rng.uniform(-25, 25, 80), np.zeros(80)], 1)
    amp = rng.standard_normal(80) + 1j * rng.standard_normal(80)
    P = len(col.ant)
    keep = np.ones(P, bool)
    for c in (P // 5, P // 2, 4 * P // 5):
        keep[c:c + rng.integers(5, 11)] = False
        keep[c + 30:c + 40:2] = False
    S = simulate_device(col, tg, amp).to_host()[keep]
    ant = col.ant[keep]
    ref = np.linalg.norm(ant, axis=1)
    from fastsar_amd import _hip
    pts = _grid_points().reshape(-1, 3)
    nfft = 1 << int(np.ceil(np.log2(64 * col.K)))
    want = bp_points_f64_device(_hip.DeviceArray.from_host(np.ascontiguousarray(S)), ant, ref, col.fmin, col.df, pts,
                                nfft, True, None, None, chunk=64).reshape(128, 128)
    print(f'dropped-pulse spotlight: {P - keep.sum()} of {P} pulses dropped, {col.K} samples', flush=True)
    return dict(col=col, S=S, ant=ant, ref=ref, want=want, tg=tg, amp=amp, keep=keep)


def test_dropped_pulses_mosaic(dropped):
    d = dropped
    fx = dict(S=d['S'], fmin=d['col'].fmin, df=d['col'].df, ref=d['ref'])
    info = []
    img = pt.form_mosaic(fx, d['ant'], ORIGIN, 128, 128, 0.5, 0.5, E1, E2, patch=(64, 64), backend='hip', info=info)
    e = rel_db(img, d['want'])
    print(f'  mosaic of {len(info)} patches: {e:.1f} dB; T {[i["T"] for i in info]}, levels {[i["levels"] for i in info]}, '
          f'gated {[i["gated"] for i in info]}, filled {[i["filled"] for i in info]}, reused {[i["reused"] for i in info]}')
    assert len(info) == 4 and all(i['filled'] > 0 for i in info)
    assert e < -45, e
    alone = pt.form_mosaic(fx, d['ant'], ORIGIN, 128, 128, 0.5, 0.5, E1, E2, patch=(64, 64), backend='hip', keep=0)
    assert any(i['reused'] for i in info)
    assert np.array_equal(alone, img)


def test_own_reference_ranges(dropped):
    """per-pulse reference ranges other than |a_p|: a scatterer adds exp(-j 4 pi f (|x - a| - ref) / c), so moving ref to
    ref' multiplies the history by exp(+j 4 pi f (ref' - ref) / c)"""
    d = dropped
    col = d['col']
    P = len(d['ant'])
    off = 3.0 * np.sin(np.linspace(0, 3 * np.pi, P)) + 1.7
    ref2 = d['ref'] + off
    f = col.fmin + col.df * np.arange(col.K)
    S2 = (d['S'] * np.exp(4j * np.pi * f[None, :] * (ref2 - d['ref'])[:, None] / C)).astype(np.complex64)
    a = pt.form_mosaic(dict(S=d['S'], fmin=col.fmin, df=col.df, ref=d['ref']), d['ant'], ORIGIN, 128, 128, 0.5, 0.5, E1, E2,
                       patch=(64, 64), backend='hip')
    b = pt.form_mosaic(dict(S=S2, fmin=col.fmin, df=col.df, ref=ref2), d['ant'], ORIGIN, 128, 128, 0.5, 0.5, E1, E2,
                       patch=(64, 64), backend='hip')
    e = rel_db(b, a)
    print(f'  own reference ranges: {e:.1f} dB')
    assert e < -60, e


def test_windowed_mosaic(dropped):
    """a beam (the along-track direction cosine, scaled so that |u| <= 1 spans the aperture about broadside) and a
    Hann window: against float64 backprojection with the same weight per pixel and pulse"""
    d = dropped
    col, ant = d['col'], d['ant']
    span = 0.9 * np.abs(ant[:, 1]).max() / np.linalg.norm(ant[0, [0, 2]])

    def beam(idx, pts):
        w = np.asarray(pts, np.float64)[None, :, :] - ant[idx][:, None, :]
        return (w[..., 1] / np.linalg.norm(w, axis=-1)) / span

    def awin(u):
        return 0.5 + 0.5 * np.cos(np.pi * u)
    fx = dict(S=d['S'], fmin=col.fmin, df=col.df, ref=d['ref'])
    img = pt.form_mosaic(fx, ant, ORIGIN, 128, 128, 0.5, 0.5, E1, E2, patch=(64, 64), backend='hip', beam=beam, awin=awin)
    pts = _grid_points(128)
    sub = pts[32:96, 32:96]                                       # the centre, where all four patches meet

    def weight(q, pul):
        return awin(np.clip(beam(np.asarray(pul), q), -1.0, 1.0))
    want = oracle.bp_points_weighted(d['S'], ant, col.fmin, col.df, sub, weight, upsample=32)
    e = rel_db(img[32:96, 32:96], want)
    print(f'  windowed mosaic (Hann over the beam): {e:.1f} dB on the central 64 x 64 pixels')
    assert e < -40, e
