"""Many images in one hipGraph launch (fastsar_amd.batch): every replay of the graph gives each image bit for bit
as its own former call does, with eager work between replays; the NaN scan of each image; the autotuner's choice
and its cache; the rejections."""
import ctypes

import numpy as np
import pytest

import oracle
from fastsar_amd import _hip, sim
from fastsar_amd.api import ImageFormer
from fastsar_amd.batch import Batch, autotune, _bind, _TUNED
from fastsar_amd.exact import simulate_device

pytestmark = pytest.mark.gpu


def _small_batch(n=6):
    scenes = sim.uneven_batch(seed=11, n=n, sides=(96, 384), pulses=(256, 1024), samples=(256, 1024))
    formers, hists = [], []
    for s in scenes:
        col, g = s['col'], s['grid']
        hists.append(simulate_device(col, s['pos'], s['amp'], clutter_sigma=s['sigma'], seed=3))
        formers.append(ImageFormer(col.ant, col.fmin, col.df, col.K, **g, T=8, levels=3, atten=80.0))
    return scenes, formers, hists


def test_batch_replays_match_single_calls():
    scenes, formers, hists = _small_batch()
    print('batch of', len(formers), [s['grid']['nx'] for s in scenes], flush=True)
    eager = [f(S) for f, S in zip(formers, hists)]
    b = Batch(formers, hists, streams=3)
    for rep in range(3):
        r = b.run()
        assert r['bad'] == [-1] * len(formers)
        assert np.all(np.diff(np.sort(r['done_ms'])) >= 0) and r['done_ms'].max() <= r['ms'] * 1.5 + 1.0
        for i, im in enumerate(b.images):
            assert np.array_equal(im.to_host(), eager[i]), (rep, i)
        # eager work between replays: a former's own call and a synchronization
        formers[rep](hists[rep])
        _hip.check(_hip.lib().fsh_synchronize())
    loop = b.run_loop()
    assert len(loop['each_ms']) == len(formers)
    for i, im in enumerate(b.images):
        assert np.array_equal(im.to_host(), eager[i])


def test_batch_reports_nonfinite_per_image():
    scenes, formers, hists = _small_batch(3)
    bad = hists[1].to_host()
    bad[5, 9] = np.nan
    hb = _hip.DeviceArray.from_host(bad)
    for f, S in zip(formers, hists):
        f(S)
    b = Batch(formers, [hists[0], hb, hists[2]], streams=2)
    r = b.run()
    assert r['bad'] == [-1, 5, -1]


def test_batch_rejections():
    scenes, formers, hists = _small_batch(2)
    with pytest.raises(ValueError, match='twice'):
        Batch([formers[0], formers[0]], [hists[0], hists[0]])
    with pytest.raises(ValueError, match='one history per former'):
        Batch(formers, hists[:1])
    # enqueue before the former has made its group buffers
    col, g = scenes[0]['col'], scenes[0]['grid']
    fresh = ImageFormer(col.ant, col.fmin, col.df, col.K, **g, T=8, levels=3)
    im = _hip.DeviceArray((g['nx'], g['ny']), np.complex64)
    L = _bind()
    assert L.fsh_ffbp_enqueue(fresh._form.h, ctypes.c_void_p(hists[0].ptr), ctypes.c_void_p(im.ptr), None) == _hip.EINVAL


def test_autotune_meets_target_and_caches():
    s = sim.uneven_batch(seed=12, n=1, sides=(256, 257), pulses=(1024, 1025), samples=(1024, 1025))[0]
    col, g = s['col'], s['grid']
    dS = simulate_device(col, s['pos'], s['amp'], clutter_sigma=s['sigma'], seed=4)
    _TUNED.clear()
    f = autotune(col.ant, col.fmin, col.df, col.K, **g, S=dS, target_db=-55.0)
    print('autotuned', f.tuned, flush=True)
    assert f.predicted_error_db <= -55.0 and f.tuned['tried'] >= 1
    f2 = autotune(col.ant, col.fmin, col.df, col.K, **g, S=dS, target_db=-55.0)
    assert f2.tuned['tried'] == 0 and (f2.T, f2.levels) == (f.T, f.levels)
    # the tuned former's image against the float64 checker on the whole (small) grid
    wp, wk = oracle.window(col.Np, col.K)
    ref = oracle.bp_points(dS.to_host() * wp[:, None] * wk[None, :], col.ant, col.fmin, col.df,
                           oracle.plane_points(**g), upsample=16, interp='cubic')
    e = oracle.rel_db(f(dS), ref)
    print(f'  error {e:.1f} dB')
    assert e < -60
