"""The hip former with aperture weights, and wide-angle apertures, against float64 exact backprojection.

Aperture weights (ImageFormer(aperture_weight=...), applied per final subaperture and tile in the final stage):
unit weights give the unweighted image; weights over pulses alone give the image of the history times those
weights (the final stage's per-subaperture mean, -43 dB in fastsar's own check, tests/test_units.py:233-245); a
weight that also varies across the scene against the oracle's backprojection with the same weight at every pixel
and pulse. Wide angle: apertures of 10 to 360 degrees around a scene, each imaged at its own resolution
(tests/test_wide_angle.py:30-55, fastsar's limit -45 dB)."""
import numpy as np
import pytest

import oracle
from fastsar_amd import ImageFormer, form_image, sim

pytestmark = pytest.mark.gpu

C = 299792458.0
GRID = dict(nx=128, ny=128, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
E1, E2 = np.array(GRID['e1']), np.array(GRID['e2'])


def _hann(P):
    return lambda p: 0.5 - 0.5 * np.cos(2 * np.pi * (np.asarray(p) + 0.5) / P)


@pytest.fixture(scope='module')
def scene():
    rng = np.random.default_rng(3)
    col = sim.make_collect(res=0.5, scene=60.0, r0=5e3)
    tg = np.stack([rng.uniform(-25, 25, 30), rng.uniform(-25, 25, 30), np.zeros(30)], 1)
    S = sim.simulate_brute(col, tg, rng.standard_normal(30) + 1j * rng.standard_normal(30)).astype(np.complex64)
    return col, S


def _former(col, K, **kw):
    return ImageFormer(col.ant, col.fmin, col.df, K, **GRID, T=16, levels=2, window=False, **kw)


def test_unit_and_pulse_weights(scene):
    col, S = scene
    P, K = S.shape
    print(f'aperture weights: {P} x {K}, 128^2', flush=True)
    plain = _former(col, K)(S)
    ones = _former(col, K, aperture_weight=lambda q, p: np.ones((len(p), len(q))))(S)
    e1 = oracle.rel_db(ones, plain)
    hann = _hann(P)
    pw = _former(col, K, aperture_weight=lambda q, p: np.repeat(hann(p)[:, None], len(q), 1))(S)
    e2 = oracle.rel_db(pw, _former(col, K)((S * hann(np.arange(P))[:, None]).astype(np.complex64)))
    print(f'  unit weights against none {e1:.1f} dB; Hann over pulses in the final stage against S x Hann {e2:.1f} dB')
    assert e1 < -100 and e2 < -43


def test_pixel_dependent_weight_against_weighted_backprojection(scene):
    """a weight that falls off across the scene along e1 and e2 and over pulses, against the float64 backprojection
    that applies it at every pixel and pulse; the same former without a weight against the unweighted
    backprojection sets the floor"""
    col, S = scene
    P, K = S.shape
    hann = _hann(P)

    def wfun(q, p):
        q = np.asarray(q)
        return hann(p)[:, None] * (1.0 + 0.004 * (q @ E1) - 0.003 * (q @ E2))[None, :]
    pts = oracle.plane_points(**GRID)
    img = _former(col, K, aperture_weight=wfun)(S)
    want = oracle.bp_points_weighted(S, col.ant, col.fmin, col.df, pts, wfun, upsample=32)
    plain = _former(col, K)(S)
    floor = oracle.rel_db(plain, oracle.bp_points(S, col.ant, col.fmin, col.df, pts, upsample=32, interp='cubic'))
    e = oracle.rel_db(img, want)
    print(f'  pixel-dependent weight {e:.1f} dB against the weighted backprojection (unweighted floor {floor:.1f} dB)')
    assert e < -45 and e < floor + 10


SPANS = (10.0, 45.0, 120.0, 360.0)


@pytest.fixture(scope='module')
def draws():
    """fastsar's wide-angle targets (tests/test_wide_angle.py:20-27): one generator, in the order of the spans"""
    rng = np.random.default_rng(4)
    out = {}
    for span in SPANS:
        u = rng.uniform(-0.35, 0.35, 10), rng.uniform(-0.35, 0.35, 10)
        out[span] = u, rng.standard_normal(10) + 1j
    return out


@pytest.mark.parametrize('span_deg', SPANS, ids=[f'{s:g}deg' for s in SPANS])
def test_wide_angle(draws, span_deg):
    """an aperture of span_deg around a scene at 1.5 km and 40 degrees grazing, X band with 1.5 GHz of band, 256^2
    pixels at 0.8 of the resolution lambda / (4 sin(span / 2)) (lambda / 4 for the circle)"""
    fc, B, r0, graze, n = 9.6e9, 1.5e9, 1500.0, np.deg2rad(40), 256
    lam = C / fc
    span = np.deg2rad(span_deg)
    sp = 0.8 * lam / (4 * np.sin(min(span, np.pi) / 2))
    side = n * sp
    df = C / (2 * 1.25 * np.sqrt(2) * side)
    K = max(16, int(B / df) // 2 * 2)
    P = int(np.ceil(span * 2 * 1.25 * np.sqrt(2) * side / lam)) // 2 * 2
    th = np.linspace(-span / 2, span / 2, P, endpoint=span_deg < 360)
    ant = np.stack([-r0 * np.cos(graze) * np.cos(th), r0 * np.cos(graze) * np.sin(th), np.full(P, r0 * np.sin(graze))], 1)
    col = sim.Collect(fmin=fc - B / 2, df=df, K=K, ant=ant, res=sp)
    (ux, uy), amp = draws[span_deg]
    tg = np.stack([ux * side, uy * side, np.zeros(10)], 1)
    S = sim.simulate_brute(col, tg, amp)
    grid = dict(nx=n, ny=n, spx=sp, spy=sp, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    print(f'aperture {span_deg:g} deg: {P} pulses x {K} samples, {n}^2 of {sp * 1e3:.1f} mm', flush=True)
    ff = form_image(S, ant, col.fmin, col.df, **grid, backend='hip')
    wp, wk = oracle.window(P, K)
    ex = oracle.bp_points(S * wp[:, None] * wk[None, :], ant, col.fmin, col.df, oracle.plane_points(**grid), upsample=16)
    e = oracle.rel_db(ff, ex)
    print(f'  ffbp against exact {e:.1f} dB')
    assert e < -45, (span_deg, e)
