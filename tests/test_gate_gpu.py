"""The range gate on the device: against the float64 oracle on identical inputs (every chunking, host and device
histories, windowed or not), and the gated former against the whole-history former on the reference's wide-band
scene (tests/test_api.py:132-179)."""
import numpy as np
import pytest

import oracle
from fastsar_amd import _hip, gate, sim
from fastsar_amd.api import ImageFormer, form_image
from fastsar_amd.exact import simulate_device

pytestmark = pytest.mark.gpu

C = sim.C


def _wide_band_collect(K, P):
    res, r0 = 0.5, 30e3
    df = C / (2 * res) / K
    y = (np.arange(P) - (P - 1) / 2) * (r0 * C / 9.6e9 / (2 * res) / P)
    psi = np.radians(40.0)
    ant = np.stack([np.full(P, -r0 * np.cos(psi)), y, np.full(P, r0 * np.sin(psi))], 1)
    return sim.Collect(fmin=9.6e9 - (K // 2) * df, df=df, K=K, ant=ant, res=res)


GRID = dict(nx=96, ny=96, spx=0.4, spy=0.4, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))


@pytest.mark.parametrize('window', [False, True])
@pytest.mark.parametrize('chunk', [None, 37])
@pytest.mark.parametrize('on_device', [False, True])
def test_gate_vs_oracle(window, chunk, on_device, monkeypatch):
    """Random histories (the gate is linear: every bin of either transform is exercised) against the complex128
    oracle; chunk 37 splits 101 pulses into chunks of 37, 37 and 27."""
    if chunk:
        monkeypatch.setenv('FSH_GATE_CHUNK', str(chunk))
    P, K = 101, 6000
    col = _wide_band_collect(K, P)
    g = gate.gate_plan(col.ant, None, col.fmin, col.df, K, **GRID)
    rng = np.random.default_rng(7)
    S = (rng.standard_normal((P, K)) + 1j * rng.standard_normal((P, K))).astype(np.complex64)
    win = oracle.window(P, K) if window else None
    print(f'gate window={window} chunk={chunk} device={on_device}: Kp {g["Kp"]} N {g["N"]} kept {g["k1"] - g["k0"]}',
          flush=True)
    G = gate.RangeGate(g, P, K, window=win)
    got = G(_hip.DeviceArray.from_host(S) if on_device else S).to_host()
    want = oracle.range_gate(S, g, win)
    e = oracle.plain_db(got, want)
    print(f'  {e:.1f} dB')
    assert e < -100


def test_gate_rejects():
    P, K = 20, 500
    col = _wide_band_collect(K, P)
    g = gate.gate_plan(col.ant, None, col.fmin, col.df, K, **GRID)
    G = gate.RangeGate(g, P, K)
    S = np.ones((P, K), np.complex64)
    S[13, 400] = np.inf
    with pytest.raises(ValueError, match='non-finite.*pulse 13'):
        G(S)
    with pytest.raises(ValueError, match='phase history must be'):
        G(S[:, :-1])
    with pytest.raises(ValueError, match='window'):
        gate.RangeGate(g, P, K, window=(np.ones(P), np.ones(K + 1)))
    bad = dict(g, k1=g['N'] + 1)
    with pytest.raises(ValueError, match='invalid gate'):
        gate.RangeGate(bad, P, K)


@pytest.fixture(scope='module')
def wide_band():
    """The reference's wide-band scene (tests/test_api.py:132-148): 24,000 samples and 384 pulses for an image of
    40 m; 5 points in the image and 4000 clutter points over 3 km of range around it, simulated on the device."""
    rng = np.random.default_rng(3)
    K, P = 24000, 384
    col = _wide_band_collect(K, P)
    pts = np.array([[0.0, 0.0, 0], [11.3, -7.1, 0], [-14.2, 12.4, 0], [15.1, 16.2, 0], [-9.3, -15.7, 0]])
    cl = np.stack([(rng.random(4000) - 0.5) * 3000, (rng.random(4000) - 0.5) * 60, np.zeros(4000)], 1)
    amp = np.concatenate([np.full(5, 30.0), (rng.standard_normal(4000) + 1j * rng.standard_normal(4000)) / np.sqrt(2)])
    dS = simulate_device(col, np.concatenate([pts, cl]), amp)
    return col, dS


@pytest.mark.parametrize('window', [True, False])
def test_range_gate_former(wide_band, window):
    """The phase history gated in range to the image's swath: fewer samples, a first level that decimates less, and
    the image of the whole history to the former's accuracy (the reference's bound, -55 dB)."""
    col, dS = wide_band
    args = (col.ant, col.fmin, col.df, col.K)
    print(f'range gate former window={window}', flush=True)
    full = ImageFormer(*args, **GRID, window=window, range_gate=False)
    gated = ImageFormer(*args, **GRID, window=window, range_gate=True)
    assert full.range_gate is None and ImageFormer(*args, **GRID).range_gate is None
    g = gated.range_gate
    assert g['K'] == col.K and g['kept'] < col.K // 20 and 'range gate' in g['note']
    assert gated._plan['levels'][0]['Dk'] < 4 < full._plan['levels'][0]['Dk']
    a, b = full(dS), gated(dS)
    e = oracle.plain_db(b, a)
    print(f'  gated to {g["kept"]} of {col.K} samples; against the whole history {e:.1f} dB')
    assert e < -55
    S = dS.to_host()
    assert np.array_equal(gated(S), b)
    assert np.array_equal(form_image(S, *args[:3], **GRID, window=window, range_gate=True), b)
    bad = S.copy()
    bad[7, 100] = np.nan
    with pytest.raises(ValueError, match='non-finite'):
        gated(bad)
    with pytest.raises(ValueError, match='phase history must be'):
        gated(S[:, :-1])
    with pytest.raises(ValueError, match='range_gate'):
        ImageFormer(*args, **GRID, range_gate='yes')
    mem = gated.memory()
    assert mem['parts']['range_gate'] > 0 and mem['parts']['history'] == 8 * col.Np * g['kept']


def test_small_scene_not_gated():
    """An image that fills the swath is not gated (tests/test_api.py:176-178)."""
    small = sim.make_collect(res=0.5, scene=60.0, r0=5e3)
    grid = dict(nx=256, ny=256, spx=0.25, spy=0.25, e1=(1.0, 0.0, 0.0), e2=(0.0, 1.0, 0.0))
    assert ImageFormer(small.ant, small.fmin, small.df, small.K, **grid, range_gate=True).range_gate is None
