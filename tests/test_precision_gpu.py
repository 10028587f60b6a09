"""16-bit histories on the device: the decode is exact (a 16-bit history forms the same image, bit for bit, as its
decoded complex64), and the precision table of the small spotlight scene (512 pulses x 1024 samples -> 512^2 px):
FFBP and exact backprojection from complex64, float16 (per-pulse power-of-two scaled) and bfloat16 storage, each
against the FP64 oracle."""
import numpy as np
import pytest

import oracle
from fastsar_amd import _hip, precision as pr, sim
from fastsar_amd.api import ImageFormer
from fastsar_amd.exact import ExactFormer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def c1():
    col, S, pos, amp, grid = sim.config1(seed=1)
    return col, S, grid


@pytest.mark.parametrize('dtype', ['float16', 'bfloat16'])
@pytest.mark.parametrize('algo', ['ffbp', 'bp'])
def test_decode_is_exact(c1, dtype, algo):
    col, S, grid = c1
    print(f'decode {algo} {dtype}', flush=True)
    mk = (lambda: ImageFormer(col.ant, col.fmin, col.df, col.K, **grid)) if algo == 'ffbp' else \
         (lambda: ExactFormer(col.ant, col.fmin, col.df, col.K, **grid))
    f = mk()
    h = pr.pack(S, dtype)
    a = f(pr.unpack(h))
    assert np.array_equal(f(h), a)
    assert np.array_equal(f(h.to_device()), a)
    bad = S.copy()
    bad[21, 7] = np.nan
    with pytest.raises(ValueError, match='pulse 21'):
        f(pr.pack(bad, dtype))


def test_precision_table(c1):
    """dB against the FP64 oracle (cubic reads of 64x oversampled float64 profiles) per storage type. The storage's
    rounding is white noise that the formation carries into the image at about its level relative to the history, so
    an image's error power is at most the complex64 image's plus twice the storage's (the round trip's error power
    relative to the history's): float16 (-74 dB here) is invisible under FFBP (-65 dB) and shows under exact
    backprojection (-80 dB); bfloat16 (-59 dB) dominates both."""
    col, S, grid = c1
    wp, wk = oracle.window(*S.shape)
    ref = oracle.bp_points(S * wp[:, None] * wk[None, :], col.ant, col.fmin, col.df, oracle.plane_points(**grid),
                           upsample=64, interp='cubic')
    rows, store = {}, {}
    for dt in ('float16', 'bfloat16'):
        back = pr.unpack(pr.pack(S, dt))
        store[dt] = np.sum(np.abs(back - S) ** 2) / np.sum(np.abs(S) ** 2)
        print(f'  storage {dt}: {10 * np.log10(store[dt]):.1f} dB')
    for algo in ('ffbp', 'bp'):
        f = ImageFormer(col.ant, col.fmin, col.df, col.K, **grid, T=8, atten=80.0) if algo == 'ffbp' else \
            ExactFormer(col.ant, col.fmin, col.df, col.K, **grid)
        for dt in ('complex64', 'float16', 'bfloat16'):
            img = f(S if dt == 'complex64' else pr.pack(S, dt))
            rows[(algo, dt)] = oracle.rel_db(img, ref)
    for (algo, dt), e in rows.items():
        print(f'  {algo:4s} {dt:9s} {e:7.1f} dB')
    for algo in ('ffbp', 'bp'):
        for dt in ('float16', 'bfloat16'):
            bound = 10 * np.log10(10 ** (rows[(algo, 'complex64')] / 10) + 2 * store[dt])
            assert rows[(algo, dt)] < bound, (algo, dt, rows[(algo, dt)], bound)
    assert rows[('ffbp', 'complex64')] < -60 and rows[('bp', 'complex64')] < -65


@pytest.mark.parametrize('kind', ['CI4', 'CI2'])
@pytest.mark.parametrize('amp', [False, True])
@pytest.mark.parametrize('algo', ['ffbp', 'bp'])
def test_cphd_integer_samples(c1, kind, amp, algo):
    """CPHD integer signal arrays converted on the device: the image equals, bit for bit, the image of the history
    the reference's reader would build on the host (float32(Re, Im) times AmpSF in float32, fastsar/io.py:160-172)"""
    col, S, grid = c1
    rng = np.random.default_rng(4)
    lim = 32000 if kind == 'CI4' else 120
    q = np.abs(np.stack([S.real, S.imag], -1)).max()
    ints = np.rint(np.stack([S.real, S.imag], -1) / q * lim)
    raw = ints.astype('>i2') if kind == 'CI4' else ints.astype(np.int8)
    a = (rng.uniform(0.5, 2.0, col.Np).astype(np.float32)) if amp else None
    host = np.empty(S.shape, np.complex64)
    host.real, host.imag = raw[..., 0], raw[..., 1]
    if a is not None:
        host *= a[:, None]
    f = ImageFormer(col.ant, col.fmin, col.df, col.K, **grid) if algo == 'ffbp' else \
        ExactFormer(col.ant, col.fmin, col.df, col.K, **grid)
    print(f'cphd {kind} amp={amp} {algo}', flush=True)
    want = f(host)
    sig = pr.CphdSignal(raw, a)
    assert np.array_equal(f(sig), want)
    assert np.array_equal(f(sig.to_device()), want)


def test_float16_intermediates(c1):
    """precision='float16': the two-pass levels' range-filtered rows stored as scaled float16 pairs. The image moves
    by the storage's rounding only (well under the float32 image's own error) and its error against the FP64 oracle
    stays within 0.5 dB"""
    col, S, grid = c1
    wp, wk = oracle.window(*S.shape)
    ref = oracle.bp_points(S * wp[:, None] * wk[None, :], col.ant, col.fmin, col.df, oracle.plane_points(**grid),
                           upsample=64, interp='cubic')
    kw = dict(T=8, levels=4, atten=80.0)
    a = ImageFormer(col.ant, col.fmin, col.df, col.K, **grid, **kw)(S)
    f = ImageFormer(col.ant, col.fmin, col.df, col.K, **grid, **kw, precision='float16')
    b = f(S)
    d, ea, eb = oracle.plain_db(b, a), oracle.rel_db(a, ref), oracle.rel_db(b, ref)
    print(f'float16 intermediates: {d:.1f} dB from float32; oracle {eb:.1f} (float32 {ea:.1f}) dB')
    assert -100 < d < -65 and eb < ea + 0.5
    # a history 2^40 times larger: the per-call scale keeps the float16 rows in range
    big = f((S * np.float32(2.0 ** 40)).astype(np.complex64))
    assert oracle.plain_db(big / np.float32(2.0 ** 40), b) < -100


def test_phase_rereference_on_device(c1):
    """former.rereference(coef): the CPHD read's phase re-reference (fastsar/io.py:384-420) applied on the device
    equals the history multiplied on the host in float64 (float32 rounding only); None removes it"""
    col, S, grid = c1
    rng = np.random.default_rng(9)
    coef = 2 * rng.uniform(-3.0, 3.0, col.Np) / sim.C           # reference moves of up to 3 m
    f = col.fmin + col.df * np.arange(col.K)
    Sr = (S * np.exp(-2j * np.pi * f[None, :] * coef[:, None])).astype(np.complex64)
    former = ImageFormer(col.ant, col.fmin, col.df, col.K, **grid)
    want = former(Sr)
    former.rereference(coef)
    got = former(S)
    e = oracle.plain_db(got, want)
    print(f're-reference on the device: {e:.1f} dB from the host-side re-reference')
    assert e < -110
    former.rereference(None)
    assert np.array_equal(former(Sr), want)
    with pytest.raises(ValueError, match='one finite value per pulse'):
        former.rereference(coef[:-1])


def test_phase_rereference_exact_former(c1):
    """ExactFormer.rereference(coef): the same re-reference on the exact former's window pass equals the history
    multiplied on the host; None removes it"""
    col, S, grid = c1
    rng = np.random.default_rng(10)
    coef = 2 * rng.uniform(-3.0, 3.0, col.Np) / sim.C
    f = col.fmin + col.df * np.arange(col.K)
    Sr = (S * np.exp(-2j * np.pi * f[None, :] * coef[:, None])).astype(np.complex64)
    former = ExactFormer(col.ant, col.fmin, col.df, col.K, **grid)
    want = former(Sr)
    former.rereference(coef)
    got = former(S)
    e = oracle.plain_db(got, want)
    print(f'exact former, re-reference on the device: {e:.1f} dB from the host-side re-reference')
    assert e < -110
    former.rereference(None)
    assert np.array_equal(former(Sr), want)
    with pytest.raises(ValueError, match='one finite value per pulse'):
        former.rereference(coef[:-1])
