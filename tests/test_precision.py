"""16-bit phase-history storage (fastsar_amd.precision), on the CPU: the float16 packing against numpy's own
float16 conversion and the per-pulse exponent rule, the bfloat16 packing against brute-force nearest-even rounding
computed in float64, and the round trip's error bound."""
import numpy as np
import pytest

from fastsar_amd import precision as pr


def _hist(seed=0, P=64, K=300):
    rng = np.random.default_rng(seed)
    scale = 10.0 ** rng.uniform(-8, 8, P)[:, None]                 # pulses over 16 decades of level
    S = (rng.standard_normal((P, K)) + 1j * rng.standard_normal((P, K))) * scale
    S[3] = 0                                                        # an all-zero pulse
    return S.astype(np.complex64)


def test_float16_exponents_and_rounding():
    S = _hist()
    h = pr.pack(S, 'float16')
    e = pr.pulse_exponents(S)
    m = np.maximum(np.abs(S.real).max(1), np.abs(S.imag).max(1)).astype(np.float64)
    top = m * 2.0 ** e
    nz = m > 0
    assert np.all((top[nz] > 2 ** 13) & (top[nz] <= 2 ** 14)) and e[3] == 0
    assert np.array_equal(h.inv, (2.0 ** -e).astype(np.float32))
    # the stored halves are numpy's float16 conversion of the exactly scaled components
    g = (2.0 ** e).astype(np.float32)[:, None]
    assert np.array_equal((h.data & 0xFFFF).astype(np.uint16), (S.real * g).astype(np.float16).view(np.uint16))
    assert np.array_equal((h.data >> 16).astype(np.uint16), (S.imag * g).astype(np.float16).view(np.uint16))
    # round trip: within half a float16 ulp at the top binade (4 in (2^13, 2^14]: under 2^-11 of the pulse's top)
    back = pr.unpack(h)
    err = np.maximum(np.abs(back.real - S.real), np.abs(back.imag - S.imag))
    assert np.all(err <= 2.0 ** -11 * m[:, None])
    rel = np.sum(np.abs(back - S) ** 2) / np.sum(np.abs(S) ** 2)
    print(f'float16 storage: {10 * np.log10(rel):.1f} dB')
    assert 10 * np.log10(rel) < -70


def _bf16_brute(x):
    """nearest bfloat16 to each float32 value, ties to the even significand, by enumeration of the two neighbours in
    float64"""
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32)
    lo = (u & np.uint32(0xFFFF0000)).view(np.float32).astype(np.float64)      # truncation toward zero
    step = np.ldexp(1.0, np.frexp(np.abs(lo) + (lo == 0))[1] - 8)             # one bfloat16 ulp at |lo|
    step = np.where(lo == 0, np.ldexp(1.0, -133), step)
    hi = np.where(np.signbit(x), lo - step, lo + step)
    xd = x.astype(np.float64)
    dl, dh = np.abs(xd - lo), np.abs(hi - xd)
    even_lo = ((u >> 16) & 1) == 0
    pick_hi = (dh < dl) | ((dh == dl) & ~even_lo)
    out = np.where(pick_hi, hi, lo).astype(np.float32)
    return (out.view(np.uint32) >> 16).astype(np.uint16)


def test_bfloat16_rounding():
    rng = np.random.default_rng(2)
    x = (rng.standard_normal(200000) * 10.0 ** rng.uniform(-30, 30, 200000)).astype(np.float32)
    # exact ties: the discarded half is 0x8000, both parities of the kept significand
    t = (rng.integers(0x3F800000, 0x40800000, 2000, dtype=np.uint32) & np.uint32(0xFFFF0000)) | np.uint32(0x8000)
    x = np.concatenate([x, t.view(np.float32), -t.view(np.float32)])
    assert np.array_equal(pr._bf16_bits(x), _bf16_brute(x))
    sp = np.array([np.inf, -np.inf, 0.0, -0.0, np.nan, 3.4e38], np.float32)
    b = pr._bf16_bits(sp)
    assert b[0] == 0x7F80 and b[1] == 0xFF80 and b[2] == 0 and b[3] == 0x8000 and b[4] == 0x7FC0 and b[5] == 0x7F80
    S = _hist(3)
    back = pr.unpack(pr.pack(S, 'bfloat16'))
    rel = np.abs(back - S) / np.maximum(np.abs(S), 1e-37)
    assert np.all(rel[np.abs(S) > 0] < 2.0 ** -8 * 1.5)


def test_history16_validation():
    S = _hist()
    with pytest.raises(ValueError, match='dtype'):
        pr.pack(S, 'float8')
    with pytest.raises(ValueError, match='per-pulse scales'):
        pr.History16(pr.pack(S, 'bfloat16').data, 'float16')
    with pytest.raises(ValueError, match='pulses, samples'):
        pr.pack(S[0], 'float16')
