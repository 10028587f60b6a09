"""The C ABI without a GPU: libfsh.so loads, exports every function include/fsh.h declares, and the calls that do no
device work (version, error text, argument validation) behave as declared."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from fastsar_amd import _hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope='module')
def lib():
    if not os.path.exists(_hip.LIBPATH):
        from fastsar_amd import build
        build.build()
    return _hip.lib()


def _declared():
    text = open(os.path.join(ROOT, 'include', 'fsh.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\bint\s+(fsh_[a-z0-9_]+)\s*\(', text)))


def test_header_declares_the_abi():
    names = _declared()
    assert names == _hip.exported_symbols()
    for must in ('fsh_ffbp_create', 'fsh_ffbp_form', 'fsh_ffbp_destroy', 'fsh_exact_create', 'fsh_exact_form',
                 'fsh_exact_destroy', 'fsh_rot_fir_k', 'fsh_fir_p', 'fsh_final_tiles', 'fsh_level_fused', 'fsh_level_mfma',
                 'fsh_last_error', 'fsh_version', 'fsh_device_info'):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.run(['nm', '-D', '--defined-only', _hip.LIBPATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r'\sT\s+(fsh_[a-z0-9_]+)$', out, flags=re.M))
    missing = [n for n in _declared() if n not in exported]
    assert not missing, missing
    for n in _declared():
        assert hasattr(lib, n)


def test_version_and_errors(lib):
    assert lib.fsh_version() == 100
    # argument validation happens before any device work
    assert lib.fsh_ffbp_create(None, None, None) == _hip.EINVAL
    assert 'null' in _hip.last_error()
    buf = ctypes.create_string_buffer(8)
    assert lib.fsh_last_error(buf, 8) == _hip.OK and len(buf.value) <= 7
    assert lib.fsh_last_error(None, 0) == _hip.EINVAL
    with pytest.raises(ValueError):
        _hip.check(_hip.EINVAL, 'x')
    with pytest.raises(MemoryError):
        _hip.check(_hip.ENOMEM, 'x')
    with pytest.raises(RuntimeError):
        _hip.check(_hip.ERUNTIME, 'x')


def test_stage_shape_validation(lib):
    """invalid shapes are refused with FSH_EINVAL before any launch (null device pointers are never touched)"""
    vp = None
    assert lib.fsh_rot_fir_k(vp, 0, 10, 10, vp, vp, 1, vp, 1, 0, 1, 10, 4.5, vp, vp) == _hip.EINVAL
    assert lib.fsh_fir_p(vp, 1, 10, 10, vp, 3, -1, 2, 5, vp, vp) == _hip.EINVAL
    assert lib.fsh_level_fused(vp, 1, 10, 20, vp, vp, 1, vp, 25, 12, 5, 8, vp, 3, 1, 2, 6, 9.5, vp, vp) == _hip.EINVAL
    assert lib.fsh_level_mfma(vp, 1, 10, 20, vp, vp, 1, vp, 25, 12, 5, 8, vp, 3, 1, 2, 6, 9.5, vp, vp) == _hip.EINVAL
    assert lib.fsh_final_tiles(vp, 1, 4, 4, 12, vp, vp, vp, 1.0, 1.0, 0.0, 0.0, 0.0, vp, vp, vp, vp, 0, 0, 12, 12, vp,
                               vp) == _hip.EINVAL


def test_plan_descriptor_layout(tmp_path):
    """the ctypes mirrors of fsh_level_desc, fsh_plan_desc, fsh_exact_desc and fsh_gate_desc have the layout a C
    compiler gives the header's structs (every field's offset and the sizes)"""
    from fastsar_amd.exact import ExactDesc
    from fastsar_amd.gate import GateDesc
    structs = [('fsh_level_desc', _hip.LevelDesc), ('fsh_plan_desc', _hip.PlanDesc), ('fsh_exact_desc', ExactDesc),
               ('fsh_gate_desc', GateDesc)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fsh.h"', 'int main(void) {']
    for cname, cls in structs:
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append('return 0; }')
    src = tmp_path / 'layout.c'
    src.write_text('\n'.join(lines))
    exe = tmp_path / 'layout'
    subprocess.run(['gcc', '-I', os.path.join(ROOT, 'include'), '-o', str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split('\n')
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, cls in structs:
        assert got[(cname, 'size')] == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)


def test_output_buffers_checked():
    """an image buffer given as out= is refused unless it is complex64 of exactly the image's shape (the library
    writes nx * ny * 8 bytes into it)"""
    from fastsar_amd import _hip
    ok = np.empty((4, 6), np.complex64)
    assert _hip.output(ok, (4, 6)) == (ok, False)
    img, dev = _hip.output(None, (4, 6))
    assert img.shape == (4, 6) and img.dtype == np.complex64 and not dev
    for bad in (np.empty((6, 4), np.complex64), np.empty((4, 6), np.complex128), np.empty((4, 12), np.complex64)[:, ::2],
                np.empty(24, np.complex64), [[0j] * 6] * 4):
        with pytest.raises(ValueError, match='out must be'):
            _hip.output(bad, (4, 6))
    ro = np.empty((4, 6), np.complex64)
    ro.flags.writeable = False
    with pytest.raises(ValueError, match='out must be'):
        _hip.output(ro, (4, 6))
