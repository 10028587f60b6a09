"""The native former's host side on the GPU: the group it keeps after an out-of-memory halving, batches captured on
it, its memory model against the bytes the library holds, the streamed host transfers against device buffers, and a
failed switch to float16 rows that leaves the former as it was."""
import ctypes

import numpy as np
import pytest

from fastsar_amd import ImageFormer, _hip, plan as pl, sim
from fastsar_amd import memory as mem
from fastsar_amd.batch import Batch
from fastsar_amd.ffbp_hip import make_ffbp_hip
from fastsar_amd.precision import pack

pytestmark = pytest.mark.gpu

GRID = dict(nx=128, ny=128, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))


@pytest.fixture(scope='module')
def scene():
    rng = np.random.default_rng(1)
    col = sim.make_collect(res=0.5, scene=60.0, r0=5e3)
    pos = np.stack([rng.uniform(-25, 25, 40), rng.uniform(-25, 25, 40), np.zeros(40)], 1)
    amp = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    return col, sim.simulate_brute(col, pos, amp).astype(np.complex64)


def test_halved_group_is_kept(scene, monkeypatch):
    """after an out-of-memory halving the former keeps the group that fitted: later calls make no buffers again, a
    batch captured on it replays; another requested group makes them again and the batch then refuses to replay"""
    from fastsar_amd.memory import MemoryWarning
    col, S = scene
    p = pl.make_plan(col, 128, 128, 0.5, 0.5, T=8, nlev=2, e1=GRID['e1'], e2=GRID['e2'])
    coll = pl.collection_arrays(p, col.ant)
    form = make_ffbp_hip(p, coll)
    forms = form.level_forms()
    want = make_ffbp_hip(p, coll)(S)
    dS = _hip.DeviceArray.from_host(S)
    form.set_group(1)
    form(S)
    form.set_group(0)
    cap = _pingpong(p, 2, forms) + 1                      # a ping-pong buffer of a group of 2 fits, of 4 does not
    assert _pingpong(p, 4, forms) > cap
    monkeypatch.setenv('FSH_FAIL_ALLOC_ABOVE', str(cap))
    with pytest.warns(MemoryWarning, match='continues with 2'):
        assert np.array_equal(form(S), want)
    gen, held = form.generation(), _hip.device_bytes()
    for _ in range(2):
        assert np.array_equal(form(S), want)
        assert form.info()['group'] == 2 and form.generation() == gen and _hip.device_bytes() == held
    b = Batch([_Wrap(form)], [dS])
    for _ in range(2):
        b.run()
        assert np.array_equal(b.images[0].to_host(), want)
    monkeypatch.delenv('FSH_FAIL_ALLOC_ABOVE')
    form.set_group(1)
    assert np.array_equal(form(S), want) and form.generation() != gen
    with pytest.raises(ValueError, match='create the batch again'):
        b.run()


def _pingpong(plan, g, forms):
    """bytes of one ping-pong buffer of a group of g first-level children: twice that plus the rotations and the final
    geometry is fastsar_amd.memory.hip_group_bytes"""
    big, parents = 0, 1
    for l, lv in enumerate(plan['levels']):
        outs = parents * (g if l == 0 else lv['C'])
        if forms[l] == 0:
            big = max(big, outs * lv['P'] * lv['Ko'])
        if lv['Dp'] > 1 or forms[l] != 0:
            big = max(big, outs * lv['Po'] * lv['Ko'])
        parents = outs
    return 8 * big


class _Wrap:
    """the attributes Batch reads from an ImageFormer, over a bare backend former"""

    def __init__(self, form):
        self._form, self._gate = form, None
        self.P, self.K = form.P, form.K

    def __call__(self, S, out=None):
        return self._form(S, out=out)


@pytest.mark.parametrize('kw', [dict(T=16, levels=2), dict(T=8, levels=3), dict(T=4, levels=4, atten=80.0),
                                dict(T=16, levels=2, precision='float16'), dict(T=8, levels=3, weighted=True)],
                         ids=['T16-2', 'T8-3', 'T4-4', 'float16', 'weights'])
def test_memory_model_matches_allocations(scene, kw):
    """memory()['needed'] (the model of fastsar_amd.memory, from the plan's shapes) against the device bytes the
    library holds once the former has formed an image from a host history: equal but for the few hundred bytes of
    the one-pass levels' polyphase tables"""
    col, S = scene
    kw = dict(kw)
    if kw.pop('weighted', False):
        kw['aperture_weight'] = lambda q, p: np.ones((len(p), len(q)))
    before = _hip.device_bytes()
    f = ImageFormer(col.ant, col.fmin, col.df, col.K, **GRID, **kw)
    f(S)
    held = _hip.device_bytes() - before
    m = f.memory()
    print(f'{kw}: model {m["needed"] / 1e6:.3f} MB, held {held / 1e6:.3f} MB, parts {m["parts"]}, group {m["group"]}')
    assert 0 <= held - m['needed'] <= 4096
    assert held == f._form.info()['device_bytes']
    assert m['full_speed'] and m['group'] == f._form.info()['group']


@pytest.mark.parametrize('stage_mb', [1, 64])
def test_streamed_host_transfers(scene, monkeypatch, stage_mb):
    """a host history uploaded in blocks through the pinned ring (1 MB buffers: many blocks, each ring buffer reused)
    and a host image downloaded in bands give the images of device buffers bit for bit, into a new array and into a
    caller's numpy array; a NaN in the last block is found; a float16 history streams the same way"""
    monkeypatch.setenv('FSH_STAGE_MB', str(stage_mb))
    rng = np.random.default_rng(5)
    col = sim.make_collect(res=0.5, scene=120.0, r0=5e3)
    pos = np.stack([rng.uniform(-50, 50, 30), rng.uniform(-50, 50, 30), np.zeros(30)], 1)
    S = sim.simulate_brute(col, pos, rng.standard_normal(30) + 1j * rng.standard_normal(30)).astype(np.complex64)
    grid = dict(GRID, nx=256, ny=192)
    f = ImageFormer(col.ant, col.fmin, col.df, col.K, **grid, T=8, levels=3)
    f._form.set_group(2)                                  # several bands of rows come down one after another
    print(f'{S.shape} history ({S.nbytes / 1e6:.1f} MB) through {stage_mb} MB buffers, groups of 2 of '
          f'{f._plan["levels"][0]["C"]}', flush=True)
    dS = _hip.DeviceArray.from_host(S)
    dimg = _hip.DeviceArray((256, 192), np.complex64)
    f(dS, out=dimg)
    want = dimg.to_host()
    assert np.array_equal(f(S), want)
    mine = np.full((256, 192), np.nan, np.complex64)
    assert f(S, out=mine) is mine and np.array_equal(mine, want)
    f(S, out=dimg)
    assert np.array_equal(dimg.to_host(), want)
    bad = S.copy()
    bad[len(S) - 3, 7] = np.inf
    with pytest.raises(ValueError, match=f'pulse {len(S) - 3}'):
        f(bad)
    assert np.array_equal(f(S), want)                     # and the former forms again after the failure
    h = pack(S, 'float16')
    dh = h.to_device()
    assert np.array_equal(f(h), f(dh))
    with pytest.raises(ValueError, match='out must be'):
        f(S, out=np.empty((192, 256), np.complex64))
    with pytest.raises(ValueError, match='out must be'):
        f(S, out=np.empty((256, 192), np.complex128))


def test_failed_precision_switch_leaves_former(scene, monkeypatch):
    """float16 rows need three small tables; when the second cannot be allocated the switch fails with MemoryError,
    holds nothing, and the former still forms its float32 image; a later switch succeeds"""
    col, S = scene
    f = ImageFormer(col.ant, col.fmin, col.df, col.K, **GRID, T=8, levels=3)
    want = f(S)
    nlev = len(f._plan['levels'])
    lib = _hip.lib()
    held = _hip.device_bytes()
    monkeypatch.setenv('FSH_FAIL_ALLOC_SIZE', str(2 * 4 * nlev))      # the per-level scales, after the maximum
    with pytest.raises(MemoryError):
        _hip.check(lib.fsh_ffbp_set_precision(f._form.h, 1), 'set_precision')
    assert _hip.device_bytes() == held
    assert np.array_equal(f(S), want)
    monkeypatch.delenv('FSH_FAIL_ALLOC_SIZE')
    _hip.check(lib.fsh_ffbp_set_precision(f._form.h, 1), 'set_precision')
    half = f(S)
    e = 10 * np.log10(np.sum(np.abs(half - want) ** 2) / np.sum(np.abs(want) ** 2))
    print(f'float16 rows against float32: {e:.1f} dB')
    assert e < -60
