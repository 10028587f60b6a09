"""Autofocus (fastsar_amd.autofocus.autofocus; the reference's fastsar/autofocus.py:124-165) through the hip former.

Scene and bars are the reference's fixture (tests/test_autofocus.py:14-48 there): seed 3, 0.6 m resolution over
0.5 m pixels, 40 points (amplitude 2-10, random phase) within +-26 m over 3000 clutter scatterers (0.15 complex
Gaussian) within +-30 m, 190 pulses x 190 samples at 5 km, 128 x 128 pixels, a former with T = 16 and 2 levels.
The phase errors put on the pulses are its three: 8 t^2, 6 t^3 + 5 t^4 - 4 t^5 (t from -1 to 1 over the aperture) and
a random walk smoothed by a 41-point Hann window and scaled to a 3 rad peak, each without its constant and linear
parts. Bars: the autofocused image within -25 dB of the error-free one and the phase estimate within 0.1 rad rms of
the error (less a line); an error-free history comes back within -25 dB with a flat estimate; PGA leaves an image of
the points alone within -30 dB."""
import numpy as np
import pytest

from fastsar_amd import ImageFormer, autofocus as af, sim
from fastsar_amd.metrics import rel_db

pytestmark = pytest.mark.gpu

GRID = dict(nx=128, ny=128, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
FORMER = dict(backend='hip', T=16, levels=2)
POINTS, CLUTTER = 40, 3000


def without_line(phase):
    """phase less its least-squares line over the pulse index (what PGA cannot see)"""
    k = np.arange(len(phase), dtype=np.float64)
    return phase - np.polyval(np.polyfit(k, phase, 1), k)


class Scene:
    """The fixture's collection, scatterers, history and phase errors (draws in the fixture's order)."""

    def __init__(self):
        rng = np.random.default_rng(3)
        self.col = sim.make_collect(res=0.6, scene=64.0, r0=5e3)
        xy_pt = [rng.uniform(-26, 26, POINTS), rng.uniform(-26, 26, POINTS)]
        xy_cl = [rng.uniform(-30, 30, CLUTTER), rng.uniform(-30, 30, CLUTTER)]
        self.pos = np.concatenate([np.column_stack(xy_pt + [np.zeros(POINTS)]), np.column_stack(xy_cl + [np.zeros(CLUTTER)])])
        a_pt = rng.uniform(2, 10, POINTS) * np.exp(2j * np.pi * rng.random(POINTS))
        a_cl = 0.15 * (rng.standard_normal(CLUTTER) + 1j * rng.standard_normal(CLUTTER))
        self.amp = np.concatenate([a_pt, a_cl])
        self.S = sim.simulate_brute(self.col, self.pos, self.amp).astype(np.complex64)
        P = self.col.Np
        t = np.linspace(-1.0, 1.0, P)
        hann = np.hanning(41)
        steps = np.cumsum(rng.standard_normal(P + 40))
        walk = without_line(np.convolve(steps, hann / hann.sum(), 'valid')[:P])
        self.errors = {'quadratic': without_line(8 * t ** 2),
                       'polynomial': without_line(6 * t ** 3 + 5 * t ** 4 - 4 * t ** 5),
                       'random walk': walk * (3.0 / np.abs(walk).max())}
        c = self.col
        self.former = ImageFormer(c.ant, c.fmin, c.df, c.K, **GRID, **FORMER)
        self.focused = self.former(self.S)

    def autofocus(self, S):
        c = self.col
        return af.autofocus(S, c.ant, c.fmin, c.df, **GRID, **FORMER)


@pytest.fixture(scope='module')
def scene():
    s = Scene()
    print(f'autofocus scene: {s.col.Np} pulses x {s.col.K} samples, {POINTS} points over {CLUTTER} clutter', flush=True)
    return s


@pytest.mark.parametrize('kind', ['quadratic', 'polynomial', 'random walk'])
def test_known_phase_error_is_removed(scene, kind):
    err = scene.errors[kind]
    blurred = (scene.S * np.exp(1j * err)[:, None]).astype(np.complex64)
    img, estimate = scene.autofocus(blurred)
    residual = np.sqrt(np.mean(without_line(estimate - err) ** 2))
    db_before = rel_db(scene.former(blurred).ravel(), scene.focused.ravel())
    db_after = rel_db(img.ravel(), scene.focused.ravel())
    print(f'{kind}: {np.abs(err).max():.1f} rad peak; image {db_before:.1f} -> {db_after:.1f} dB; '
          f'residual {residual:.3f} rad rms')
    assert db_before > -10, 'the error must blur the image for the test to mean anything'
    assert db_after < -25
    assert residual < 0.1


def test_error_free_history_comes_back(scene):
    img, estimate = scene.autofocus(scene.S)
    assert rel_db(img.ravel(), scene.focused.ravel()) < -25
    assert np.sqrt(np.mean(without_line(estimate) ** 2)) < 0.1


def test_pga_keeps_a_focused_image(scene):
    """PGA itself (the polar-format deramped image in, the corrected image out) on the image of the points alone"""
    c = scene.col
    deramp = np.exp(-1j * af.deramp_phase(c.ant, c.fmin, c.df, c.K, **GRID))
    pts = sim.simulate_brute(c, scene.pos[:POINTS], scene.amp[:POINTS]).astype(np.complex64)
    img = scene.former(pts) * deramp
    out, _, _ = af.pga(img)
    assert rel_db(out.ravel(), img.ravel()) < -30
    full = scene.focused * deramp
    out, _, _ = af.pga(full)
    assert rel_db(out.ravel(), full.ravel()) < -25


def test_autofocus_refuses_bad_arguments(scene):
    c = scene.col
    with pytest.raises(ValueError):
        af.autofocus(scene.S, c.ant, c.fmin, c.df, **GRID, algorithm='exact', backend='hip')
    with pytest.raises(ValueError):
        af.autofocus(scene.S[0], c.ant, c.fmin, c.df, **GRID, **FORMER)
