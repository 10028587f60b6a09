"""ctypes binding of libfsh.so (include/fsh.h): status codes mapped to Python exceptions, the plan descriptors,
and a small device-buffer type for stage-level callers.

The library is loaded from fastsar_amd/_lib/ (built in-tree by fastsar_amd.build). There is no fallback: a
missing or unloadable library raises, so no caller can silently run on another path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, '_lib', 'libfsh.so')

OK, ENOMEM, EINVAL, ENONFINITE, ERUNTIME = 0, 1, 2, 3, 4

_lib = None
c_int, c_double, c_size_t, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p
P_f32 = ctypes.POINTER(ctypes.c_float)
P_f64 = ctypes.POINTER(ctypes.c_double)
P_i32 = ctypes.POINTER(ctypes.c_int)


class LevelDesc(ctypes.Structure):
    _fields_ = [(n, c_int) for n in ('sx', 'sy', 'C', 'Dk', 'Dp', 'K', 'P', 'Ko', 'Po', 'Lk', 'plk', 'Lp', 'plp', 'nparents')] + \
               [(n, c_double) for n in ('kc', 'k0c', 'k1')] + \
               [('taps_k', P_f32), ('taps_p', P_f32), ('c0', P_f32), ('sl', P_f32),
                ('u', P_f64), ('r0', P_f64), ('d', P_f64), ('ref', P_f64)]


class PlanDesc(ctypes.Structure):
    _fields_ = [(n, c_int) for n in ('nlev', 'T', 'nx', 'ny', 'Nx', 'Ny', 'ox', 'oy', 'Pf', 'Qf', 'ntiles', 'P', 'K')] + \
               [(n, c_double) for n in ('spx', 'spy', 'a0', 'a1', 'fc2')] + \
               [('e1', c_double * 3), ('e2', c_double * 3),
                ('fu', P_f64), ('fr0', P_f64), ('fcen', P_f64), ('tile_xy', P_i32), ('wp', P_f32), ('wk', P_f32),
                ('w0', P_f32), ('gx', P_f32), ('gyr', P_f32)]


def lib():
    """The loaded library (raises when it is not built or cannot load)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f'the HIP library is not built ({LIBPATH} missing): run python -m fastsar_amd.build')
    L = ctypes.CDLL(LIBPATH)
    vp = c_void_p
    sig = {
        'fsh_version': ([], c_int),
        'fsh_device_info': ([ctypes.c_char_p, c_int], c_int),
        'fsh_last_error': ([ctypes.c_char_p, c_int], c_int),
        'fsh_malloc': ([ctypes.POINTER(vp), c_size_t], c_int),
        'fsh_free': ([vp], c_int),
        'fsh_memcpy_h2d': ([vp, vp, c_size_t], c_int),
        'fsh_memcpy_d2h': ([vp, vp, c_size_t], c_int),
        'fsh_memcpy_d2d': ([vp, vp, c_size_t], c_int),
        'fsh_memset': ([vp, c_int, c_size_t], c_int),
        'fsh_synchronize': ([], c_int),
        'fsh_copy': ([vp, vp, c_size_t, vp], c_int),
        'fsh_device_bytes': ([ctypes.POINTER(c_size_t)], c_int),
        'fsh_mem_info': ([ctypes.POINTER(c_size_t), ctypes.POINTER(c_size_t)], c_int),
        'fsh_ffbp_level_forms': ([vp, P_i32, c_int], c_int),
        'fsh_ffbp_generation': ([vp, ctypes.POINTER(ctypes.c_ulonglong)], c_int),
        'fsh_ffbp_adopt': ([vp, vp, P_i32], c_int),
        'fsh_rot_fir_k': ([vp, c_int, c_int, c_int, vp, vp, c_int, vp, c_int, c_int, c_int, c_int, c_double, vp, vp], c_int),
        'fsh_fir_p': ([vp, c_int, c_int, c_int, vp, c_int, c_int, c_int, c_int, vp, vp], c_int),
        'fsh_level_fused': ([vp, c_int, c_int, c_int, vp, vp, c_int, vp, c_int, c_int, c_int, c_int, vp, c_int, c_int, c_int,
                             c_int, c_double, vp, vp], c_int),
        'fsh_level_mfma': ([vp, c_int, c_int, c_int, vp, vp, c_int, vp, c_int, c_int, c_int, c_int, vp, c_int, c_int, c_int,
                            c_int, c_double, vp, vp], c_int),
        'fsh_level_mfma_noalloc': ([vp, c_int, c_int, c_int, vp, vp, c_int, vp, c_int, c_int, c_int, c_int, vp, c_int, c_int,
                                    c_int, c_int, c_double, vp, vp], c_int),
        'fsh_device_phases': ([vp, vp, vp, vp, c_int, c_int, c_int, c_double, c_double, vp, vp, vp], c_int),
        'fsh_final_geom': ([vp, vp, vp, c_int, c_int, P_f64, P_f64, vp, vp, vp, vp], c_int),
        'fsh_final_tiles': ([vp, c_int, c_int, c_int, c_int, vp, vp, vp, c_double, c_double, c_double, c_double, c_double,
                             vp, vp, vp, vp, c_int, c_int, c_int, c_int, vp, vp], c_int),
        'fsh_ffbp_create': ([ctypes.POINTER(PlanDesc), ctypes.POINTER(LevelDesc), ctypes.POINTER(vp)], c_int),
        'fsh_ffbp_form': ([vp, vp, c_int, vp, c_int, P_i32], c_int),
        'fsh_ffbp_form_typed': ([vp, vp, c_int, vp, c_int, vp, c_int, P_i32], c_int),
        'fsh_ffbp_set_precision': ([vp, c_int], c_int),
        'fsh_ffbp_set_ramp': ([vp, vp], c_int),
        'fsh_ffbp_set_group': ([vp, c_int], c_int),
        'fsh_ffbp_info': ([vp, P_i32, ctypes.POINTER(c_size_t)], c_int),
        'fsh_ffbp_stage_ms': ([vp, P_f32, c_int], c_int),
        'fsh_ffbp_set_profile': ([vp, c_int], c_int),
        'fsh_ffbp_profile': ([vp, P_i32, P_i32, P_f32, c_int, P_i32], c_int),
        'fsh_ffbp_destroy': ([vp], c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def exported_symbols():
    """The fsh_* names include/fsh.h declares."""
    import re
    with open(os.path.join(os.path.dirname(HERE), 'include', 'fsh.h')) as fh:
        return sorted(set(re.findall(r'\b(fsh_[a-z0-9_]+)\s*\(', fh.read())))


def last_error():
    buf = ctypes.create_string_buffer(1024)
    lib().fsh_last_error(buf, 1024)
    return buf.value.decode(errors='replace')


def check(status, what=''):
    """Map a status to an exception: ENOMEM -> MemoryError, EINVAL and ENONFINITE -> ValueError, other -> RuntimeError."""
    if status == OK:
        return
    msg = (what + ': ' if what else '') + last_error()
    if status == ENOMEM:
        raise MemoryError(msg)
    if status in (EINVAL, ENONFINITE):
        raise ValueError(msg)
    raise RuntimeError(msg)


def device_info():
    buf = ctypes.create_string_buffer(512)
    check(lib().fsh_device_info(buf, 512), 'device_info')
    name, arch, cu, lds, mem = buf.value.decode().split('|')
    return dict(name=name, arch=arch, cu=int(cu.split('=')[1]), lds=int(lds.split('=')[1]), mem=int(mem.split('=')[1]))


def available():
    """Whether the library loads and a HIP device answers."""
    try:
        device_info()
        return True
    except Exception:
        return False


def device_bytes():
    """Device bytes the library holds now (every allocation it made and has not freed)."""
    n = c_size_t()
    check(lib().fsh_device_bytes(ctypes.byref(n)), 'device_bytes')
    return n.value


def mem_info():
    """(free, total) device bytes."""
    fr, to = c_size_t(), c_size_t()
    check(lib().fsh_mem_info(ctypes.byref(fr), ctypes.byref(to)), 'mem_info')
    return fr.value, to.value


def output(out, shape):
    """(buffer, on_device) for an image of `shape`: a new host array when out is None, else out itself, which must be
    a complex64 DeviceArray or a writable C-contiguous complex64 numpy array of exactly that shape (ValueError
    otherwise: the native side writes shape[0] * shape[1] * 8 bytes into it)."""
    shape = tuple(int(n) for n in shape)
    if out is None:
        return np.empty(shape, np.complex64), False
    if isinstance(out, DeviceArray):
        if out.shape != shape or out.dtype != np.complex64:
            raise ValueError(f'out must be complex64 {shape}, got a DeviceArray {out.dtype} {out.shape}')
        return out, True
    if isinstance(out, np.ndarray):
        if out.shape != shape or out.dtype != np.complex64 or not out.flags.c_contiguous or not out.flags.writeable:
            raise ValueError(f'out must be a writable C-contiguous complex64 array of shape {shape}, got {out.dtype} '
                             f'{out.shape}')
        return out, False
    raise ValueError(f'out must be a complex64 DeviceArray or numpy array of shape {shape}, got {type(out).__name__}')


def ptr(a):
    """The data pointer of a host array or a DeviceArray, as c_void_p (None for None)."""
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return c_void_p(a.ptr)
    return c_void_p(a.ctypes.data)


class DeviceArray:
    """Device memory of a given shape and dtype, freed with the object. to_host() copies back; from_host uploads."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if np.ndim(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = c_void_p()
        check(lib().fsh_malloc(ctypes.byref(p), max(self.nbytes, 1)), 'malloc')
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(lib().fsh_memcpy_h2d(c_void_p(d.ptr), c_void_p(a.ctypes.data), a.nbytes), 'h2d')
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().fsh_memcpy_d2h(c_void_p(out.ctypes.data), c_void_p(self.ptr), self.nbytes), 'd2h')
        return out

    def zero(self):
        check(lib().fsh_memset(c_void_p(self.ptr), 0, self.nbytes), 'memset')
        return self

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.fsh_free(c_void_p(self.ptr))
        except Exception:
            pass
        self.ptr = None
