"""Phase histories stored in 16 bits per component, for the precision variants of the former.

A complex sample is one 32-bit word: Re in the low half, Im in the high half, each either

  * float16, with every pulse stored times a power of two 2^e_p chosen so that its largest component lies in
    (2^13, 2^14] (float16's precision where the pulse's energy is, and headroom below its 65504 limit); the former
    multiplies by inv[p] = 2^-e_p when it decodes, which is exact; or
  * bfloat16 (float32's exponent range, 8 significant bits), with no scale.

Both round to nearest, ties to even. The former decodes on the device before the window, so a 16-bit history costs
half the upload and half the device footprint of a complex64 one:

    h = pack(S, 'float16')           # or 'bfloat16'; History16 on the host
    img = former(h)                  # ImageFormer or ExactFormer
    dh = h.to_device()               # kept on the device between calls
"""
import numpy as np

from . import _hip

DTYPES = {'float16': 1, 'bfloat16': 2}          # FSH_C32F16, FSH_C32BF16
F16_TOP = 14                                     # the largest component of a pulse lands in (2^13, 2^14]


def _bf16_bits(x):
    """float32 -> bfloat16 bits (uint16), rounded to nearest, ties to even; NaN stays NaN."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32)
    r = ((u + np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))) >> np.uint32(16)).astype(np.uint16)
    nan = np.isnan(np.asarray(x, np.float32))
    r[nan] = np.uint16(0x7FC0)
    return r


def pulse_exponents(S):
    """e_p per pulse: 2^e_p times the pulse's largest component lies in (2^13, 2^14]; 0 for an all-zero or
    non-finite pulse (whose NaN or inf then reaches the former's scan)."""
    S = np.asarray(S)
    m = np.maximum(np.abs(S.real).max(axis=1), np.abs(S.imag).max(axis=1)).astype(np.float64)
    e = np.zeros(m.shape, np.int32)
    ok = np.isfinite(m) & (m > 0)
    e[ok] = F16_TOP - np.ceil(np.log2(m[ok])).astype(np.int32)
    return np.clip(e, -126, 126)


class TypedHistory:
    """A [P, K] history in a stored type the formers decode on the device (fsh_ffbp_form_typed): data (host ndarray
    or DeviceArray) whose first two dimensions are [P, K], code (the FSH_* type), inv float32 [P] (per-pulse factors,
    on the same side as data) or None."""

    def __init__(self, data, code, inv=None):
        self.data, self.code, self.inv = data, int(code), inv
        self.shape = tuple(data.shape[:2])
        self.on_device = isinstance(data, _hip.DeviceArray)

    def _like(self, data, inv):
        return TypedHistory(data, self.code, inv)

    def to_device(self):
        if self.on_device:
            return self
        return self._like(_hip.DeviceArray.from_host(np.ascontiguousarray(self.data)),
                          None if self.inv is None else _hip.DeviceArray.from_host(np.ascontiguousarray(self.inv, np.float32)))

    def to_host(self):
        if not self.on_device:
            return self
        return self._like(self.data.to_host(), None if self.inv is None else self.inv.to_host())


class History16(TypedHistory):
    """A [P, K] complex history in 16-bit pairs: data uint32 [P, K], dtype 'float16' or 'bfloat16', inv float32 [P]
    (float16: the per-pulse inverse scales) or None."""

    def __init__(self, data, dtype, inv=None):
        if dtype not in DTYPES:
            raise ValueError(f"dtype must be 'float16' or 'bfloat16', got {dtype!r}")
        if (inv is None) != (dtype == 'bfloat16'):
            raise ValueError('a float16 history carries per-pulse scales; a bfloat16 one carries none')
        super().__init__(data, DTYPES[dtype], inv)
        self.dtype = dtype

    def _like(self, data, inv):
        return History16(data, self.dtype, inv)


class CphdSignal(TypedHistory):
    """A CPHD signal array as it lies in the file, for the formers to convert on the device (the conversion of
    fastsar/io.py:128-172): samples [P, K, 2] big-endian int16 (CI4, '>i2') or int8 (CI2), and the vectors' AmpSF
    [P] (or None for 1). The decoded sample is float32(Re, Im) times AmpSF in float32."""

    def __init__(self, samples, amp=None):
        if isinstance(samples, _hip.DeviceArray):
            raise ValueError('CphdSignal takes the host array; call to_device() on it')
        a = np.asarray(samples)
        if a.ndim != 3 or a.shape[2] != 2:
            raise ValueError(f'CPHD samples must be [pulses, samples, 2], got shape {a.shape}')
        if a.dtype == np.dtype('>i2'):
            code = 3
        elif a.dtype == np.int8:
            code = 4
        else:
            raise ValueError(f"CPHD samples must be big-endian int16 ('>i2', CI4) or int8 (CI2), got {a.dtype}")
        inv = None
        if amp is not None:
            inv = np.ascontiguousarray(amp, np.float32)
            if inv.shape != (a.shape[0],):
                raise ValueError(f'AmpSF must hold one factor per pulse ({a.shape[0]}), got shape {inv.shape}')
        super().__init__(np.ascontiguousarray(a), code, inv)

    def _like(self, data, inv):
        t = TypedHistory(data, self.code, inv)
        return t


def pack(S, dtype='float16'):
    """S [P, K] complex -> History16 on the host."""
    S = np.asarray(S)
    if S.ndim != 2:
        raise ValueError(f'phase history must be [pulses, samples], got shape {S.shape}')
    S = S.astype(np.complex64, copy=False)
    if dtype == 'float16':
        e = pulse_exponents(S)
        g = np.ldexp(np.float32(1), e).astype(np.float32)[:, None]
        re = (S.real * g).astype(np.float16).view(np.uint16)
        im = (S.imag * g).astype(np.float16).view(np.uint16)
        inv = np.ldexp(np.float32(1), -e).astype(np.float32)
    elif dtype == 'bfloat16':
        re, im, inv = _bf16_bits(S.real), _bf16_bits(S.imag), None
    else:
        raise ValueError(f"dtype must be 'float16' or 'bfloat16', got {dtype!r}")
    data = re.astype(np.uint32) | (im.astype(np.uint32) << np.uint32(16))
    return History16(np.ascontiguousarray(data), dtype, inv)


def unpack(h):
    """History16 -> complex64 [P, K] on the host (the decode the former runs on the device)."""
    h = h.to_host()
    lo = (h.data & np.uint32(0xFFFF)).astype(np.uint16)
    hi = (h.data >> np.uint32(16)).astype(np.uint16)
    if h.dtype == 'float16':
        g = h.inv[:, None]
        re = lo.view(np.float16).astype(np.float32) * g
        im = hi.view(np.float16).astype(np.float32) * g
    else:
        re = (lo.astype(np.uint32) << np.uint32(16)).view(np.float32)
        im = (hi.astype(np.uint32) << np.uint32(16)).view(np.float32)
    return (re + 1j * im).astype(np.complex64)
