"""Autofocus of a phase history through the hip former (fastsar_amd.autofocus.autofocus; fastsar/autofocus.py:124-165),
on the reference's test scene (tests/test_autofocus.py:14-48 of the reference): 40 points over 3000 clutter
scatterers, 190 pulses x 190 samples, 128 x 128 pixels of 0.5 m, a former with T = 16 and 2 levels. A known phase
error per pulse (quadratic, higher-order polynomial, low-pass random walk) is put on the history; after autofocus
the image must match the image of the error-free history within -25 dB and the estimated phase must match the
error within 0.1 rad rms, less constant and linear terms (the reference's bars)."""
from types import SimpleNamespace

import numpy as np
import pytest

from fastsar_amd import sim, autofocus as af

pytestmark = pytest.mark.gpu

GRID = dict(nx=128, ny=128, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
FK = dict(backend='hip', T=16, levels=2)
N_PT, N_CL = 40, 3000


def rel_db(a, b):
    g = np.vdot(a, b) / np.vdot(a, a)
    return 10 * np.log10(np.sum(np.abs(g * a - b) ** 2) / np.sum(np.abs(b) ** 2))


@pytest.fixture(scope='module')
def scene():
    from fastsar_amd import ImageFormer
    rng = np.random.default_rng(3)
    col = sim.make_collect(res=0.6, scene=64.0, r0=5e3)
    pos = np.concatenate([np.stack([rng.uniform(-26, 26, N_PT), rng.uniform(-26, 26, N_PT), np.zeros(N_PT)], 1),
                          np.stack([rng.uniform(-30, 30, N_CL), rng.uniform(-30, 30, N_CL), np.zeros(N_CL)], 1)])
    amp = np.concatenate([rng.uniform(2, 10, N_PT) * np.exp(2j * np.pi * rng.random(N_PT)),
                          0.15 * (rng.standard_normal(N_CL) + 1j * rng.standard_normal(N_CL))])
    S = sim.simulate_brute(col, pos, amp).astype(np.complex64)
    P = col.Np
    p = np.arange(P, dtype=np.float64)
    t = 2 * p / (P - 1) - 1
    walk = np.convolve(np.cumsum(rng.standard_normal(P + 40)), np.hanning(41) / np.hanning(41).sum(), 'valid')[:P]
    errs = {'quadratic': 8 * t ** 2, 'polynomial': 6 * t ** 3 + 5 * t ** 4 - 4 * t ** 5, 'random walk': walk}
    former = ImageFormer(col.ant, col.fmin, col.df, col.K, **GRID, **FK)
    print(f'autofocus scene: {P} pulses, {col.K} samples, {N_PT} points over {N_CL} clutter', flush=True)
    return SimpleNamespace(col=col, pos=pos, amp=amp, S=S, p=p, errs=errs, former=former, ref=former(S))


@pytest.mark.parametrize('name', ['quadratic', 'polynomial', 'random walk'])
def test_phase_error_removed(scene, name):
    c, S, former, ref = scene.col, scene.S, scene.former, scene.ref
    e = af._remove_line(np.array(scene.errs[name]), scene.p)
    if name == 'random walk':
        e *= 3 / np.abs(e).max()
    Se = S * np.exp(1j * e)[:, None].astype(np.complex64)
    img, phi = af.autofocus(Se, c.ant, c.fmin, c.df, **GRID, **FK)
    res = af._remove_line(phi - e, scene.p)
    before, after = rel_db(former(Se), ref), rel_db(img, ref)
    print(f'{name}: peak {np.abs(e).max():.1f} rad: image {before:.1f} dB -> {after:.1f} dB, residual {res.std():.3f} rad')
    assert before > -10
    assert res.std() < 0.1 and after < -25, (name, res.std(), after)


def test_focused_image_unchanged(scene):
    c, S, former, ref = scene.col, scene.S, scene.former, scene.ref
    ramp = np.exp(-1j * af.deramp_phase(c.ant, c.fmin, c.df, c.K, **GRID))
    img, _, hist = af.pga(ref * ramp)
    img2, phi = af.autofocus(S, c.ant, c.fmin, c.df, **GRID, **FK)
    assert rel_db(img, ref * ramp) < -25 and rel_db(img2, ref) < -25
    assert af._remove_line(phi, scene.p).std() < 0.1
    pts = former(sim.simulate_brute(c, scene.pos[:N_PT], scene.amp[:N_PT]).astype(np.complex64)) * ramp
    img, _, _ = af.pga(pts)
    assert rel_db(img, pts) < -30


def test_autofocus_arguments(scene):
    c = scene.col
    with pytest.raises(ValueError):
        af.autofocus(scene.S, c.ant, c.fmin, c.df, **GRID, algorithm='exact', backend='hip')
    with pytest.raises(ValueError):
        af.autofocus(scene.S[0], c.ant, c.fmin, c.df, **GRID, **FK)
