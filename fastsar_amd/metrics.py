"""Image error in dB and the benchmark's algorithmic operation counts.

rel_db is the reference's error metric: 10 log10(sum |g a - b|^2 / sum |b|^2) with one fitted complex gain
g = <a, b> / <a, a> (tests/test_exact.py:20-22, docs/precision.md:16-17).
"""
import numpy as np


def rel_db(a, b):
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    g = np.vdot(a, b) / np.vdot(a, a)
    return float(10 * np.log10(np.sum(np.abs(g * a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def level_counts(plan):
    """Per level: parents over the whole image N, children C and the shapes, as (N, lv)."""
    out, N = [], 1
    for lv in plan['levels']:
        out.append((N, lv))
        N *= lv['C']
    return out


def stage_flops(plan):
    """Algorithmic float operations of each (stage, level) over one image:
      rotation + range FIR  6 N C P K + 4 N C P Ko Lk      (one complex multiply per sample per child; a real tap
                                                           times a complex sample per output and tap)
      pulse FIR             4 N C Po Ko Lp
      final stage           8 T^2 Pf Qf per tile (a real 2T x PfQf x 2T product)
    and ('level_fused', l) / ('level_mfma', l), the level's rotation and both filters in one pass (on VALU / with the
    range filter on the matrix cores), their sum."""
    out = {}
    for l, (N, lv) in enumerate(level_counts(plan)):
        rot = 6.0 * N * lv['C'] * lv['P'] * lv['K']
        fk = 4.0 * N * lv['C'] * lv['P'] * lv['Ko'] * lv['fir_k']['L'] if lv['Dk'] > 1 else 0.0
        out[('rot_fir_k', l)] = rot + fk
        if lv['Dp'] > 1:
            out[('fir_p', l)] = 4.0 * N * lv['C'] * lv['Po'] * lv['Ko'] * lv['fir_p']['L']
        out[('level_fused', l)] = out[('rot_fir_k', l)] + out.get(('fir_p', l), 0.0)
        out[('level_mfma', l)] = out[('level_fused', l)]
    f = plan['final']
    ntiles = f['cen'].shape[0]
    out[('final_tiles', len(plan['levels']))] = 8.0 * plan['T'] ** 2 * f['P'] * f['K'] * ntiles
    return out


def min_bytes(plan):
    """Fused-minimum HBM bytes of one image in complex64: the history once, every level's output written and read
    once, the image written once."""
    b = 8.0 * plan['levels'][0]['P'] * plan['levels'][0]['K'] + 8.0 * plan['nx'] * plan['ny']
    for N, lv in level_counts(plan):
        b += 2 * 8.0 * N * lv['C'] * lv['Po'] * lv['Ko']
    return b


def stage_bytes(plan):
    """Algorithmic HBM bytes of each (stage, level) over one image, complex64 samples and float32 rotations, each
    array touched once:
      rotation + range FIR  8 N P K (parent rows) + 8 N C P (c0, sl) + 8 N C P Ko (output)
      pulse FIR             8 N C P Ko (input) + 8 N C Po Ko (output)
      final stage           8 Pf Qf per tile (input) + 16 Pf per tile (ux, uy) + 8 nx ny (image)
      fused level           8 N P K + 8 N C P + 8 N C Po Ko (parent rows, rotations, output)."""
    out = {}
    for l, (N, lv) in enumerate(level_counts(plan)):
        out[('rot_fir_k', l)] = 8.0 * N * (lv['P'] * lv['K'] + lv['C'] * lv['P'] + lv['C'] * lv['P'] * lv['Ko'])
        if lv['Dp'] > 1:
            out[('fir_p', l)] = 8.0 * N * lv['C'] * (lv['P'] + lv['Po']) * lv['Ko']
        out[('level_fused', l)] = 8.0 * N * (lv['P'] * lv['K'] + lv['C'] * lv['P'] + lv['C'] * lv['Po'] * lv['Ko'])
        out[('level_mfma', l)] = out[('level_fused', l)]
    f = plan['final']
    ntiles = f['cen'].shape[0]
    out[('final_tiles', len(plan['levels']))] = (8.0 * f['P'] * f['K'] + 16.0 * f['P']) * ntiles + 8.0 * plan['nx'] * plan['ny']
    return out
