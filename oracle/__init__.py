"""The oracle: test infrastructure only, the checker the HIP path is compared with; never part of the product.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this package. It holds

  * a float64 C restatement of the hot path's stages (oracle/oracle.c, built by oracle/Makefile):
      rot_fir_k, fir_p, final_tiles  (fastsar/ffbp_cpu.cpp:147-469)
      bp_points  FP64 exact backprojection from float64 profiles (fastsar/bp.py:55-63, 99-138; tests/test_bp.py:23-35)
      exact_tiles  ExactFormer's tiled algorithm in float64 (fastsar/exact.py:240-328)
      range_gate   the range gate of a wide-band history in complex128 (fastsar/api.py:546-563)
      tapnet_*, ls_loss_grad  the learned decimators' network and least-squares loss in float64 (no reference
                   counterpart: the checker of the HIP training kernels)
  * the reference's own CPU kernels, compiled in place from its sources into oracle/_ref/ (ref_* below);
  * the orchestration of the CPU former restated in Python (fastsar/ffbp_cpu.py:158-291), which drives either
    set of stages: ffbp_image(..., stages='ref') is the reference's CPU path, stages='oracle' the float64 one;
  * rel_db, the reference's error metric with a fitted complex gain (tests/test_exact.py:20-22).

How each function is pinned (tests/test_oracle.py):
  rot_fir_k, fir_p, final_tiles   against the reference's compiled kernels on identical inputs (float32 rounding
                                  only, below -120 dB); final_tiles also against a brute-force double sum over
                                  (p, q) for every pixel on a tiny tile
  bp_points                       against the direct interpolation-free sum over pulses and samples on tiny cases,
                                  against a numpy complex128 backprojection (tests/test_bp.py:23-35), the error
                                  falling 12 dB per doubling of oversampling, and range periodicity
  exact_tiles                     against bp_points (cubic, -65 dB at 20 km and 600 km, tests/test_exact.py:40-50)
  ffbp_image                      stages='ref' against stages='oracle' (-100 dB), and against bp_points within the
                                  final-stage model's predicted error (tests/test_accuracy.py:29-54)
  window                          scipy.signal.windows.taylor, the library routine the reference calls
  tapnet_forward/backward,        central finite differences of the float64 forward pass and loss (relative 1e-6;
  ls_loss_grad                    tests/test_learned.py), and the loss against the direct sum over the pairs
  range_gate                      the image it forms: float64 exact backprojection of the gated and the whole history
                                  agree to -98 dB (no fitted gain), and planted mistakes in the range cut, the
                                  transform directions or the Kp / N scale each fail that check
"""
import ctypes
import os
import subprocess

import numpy as np

C = 299792458.0
HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None


def build():
    """Compile liboracle.so (and oracle/_ref/ where the reference sources are present)."""
    subprocess.run(['make', '-s', '-C', HERE], check=True)


def _load(path):
    return ctypes.CDLL(path)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, 'liboracle.so')
        if not os.path.exists(path):
            build()
        L = _load(path)
        f64 = np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS')
        i32 = np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS')
        ci, cd, cl, vp = ctypes.c_int, ctypes.c_double, ctypes.c_long, ctypes.c_void_p
        L.orc_rot_fir_k.argtypes = [f64, f64, ci, ci, ci, f64, f64, ci, f64, ci, ci, ci, ci, cd, f64, f64]
        L.orc_fir_p.argtypes = [f64, f64, ci, ci, ci, f64, ci, ci, ci, ci, f64, f64]
        L.orc_final_tiles.argtypes = [f64, f64, ci, ci, ci, ci, f64, f64, f64, f64, cd, cd, f64, f64, f64, cd, vp, vp, vp, f64, f64]
        L.orc_bp_points.argtypes = [f64, f64, ci, ci, f64, f64, f64, cl, cd, cd, ci, f64, f64]
        L.orc_exact_tiles.argtypes = [f64, f64, ci, ci, i32, f64, f64, f64, ci, ci, ci, ci, ci, ci, cd, cd, f64, f64, cd, cd, ci,
                                      f64, f64]
        _LIB = L
    return _LIB


def _avx512():
    try:
        with open('/proc/cpuinfo') as fh:
            return ' avx512f ' in fh.read().replace('\n', ' ')
    except OSError:
        return False


def have_ref():
    return any(os.path.exists(os.path.join(HERE, '_ref', f'libffbp_cpu_{v}.so')) for v in ('v4', 'v3'))


def ref_lib():
    """The reference's CPU kernels (fastsar/ffbp_cpu.cpp), compiled from its sources: the AVX-512 build on hosts
    that have AVX-512, else the AVX2 one. Their ctypes signatures follow fastsar/ffbp_cpu.py:38-46."""
    global _REF
    if _REF is None:
        v = 'v4' if _avx512() and os.path.exists(os.path.join(HERE, '_ref', 'libffbp_cpu_v4.so')) else 'v3'
        L = _load(os.path.join(HERE, '_ref', f'libffbp_cpu_{v}.so'))
        f32 = np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS')
        f64 = np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS')
        i, dbl, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        L.rot_fir_k.argtypes = [f32, f32, i, i, i, f32, f32, i, f32, i, i, i, i, dbl, f32, f32]
        L.rot_fir_k_cplx.argtypes = [f32, i, i, i, f32, f32, i, f32, i, i, i, i, dbl, f32, f32]
        L.fir_p.argtypes = [f32, f32, i, i, i, f32, i, i, i, i, f32, f32]
        L.final_tiles.argtypes = [f32, f32, i, i, i, i, f64, f64, f64, f64, dbl, dbl, f64, f64, f64, dbl, f32, f32, vp, vp, vp]
        L.ffbp_cpu_threads.restype = i
        L.isa = v
        _REF = L
    return _REF


# ------------------------------------------------------------------------------------------------ metric, window

def rel_db(a, b):
    """10 log10(sum |g a - b|^2 / sum |b|^2) with the fitted complex gain g = <a, b> / <a, a>.
    (tests/test_exact.py:20-22)"""
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    g = np.vdot(a, b) / np.vdot(a, a)
    return float(10 * np.log10(np.sum(np.abs(g * a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def plain_db(a, b):
    """10 log10(sum |a - b|^2 / sum |b|^2), no fitted gain (tests/test_api.py:18-19)."""
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(10 * np.log10(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def window(P, K, sll=35.0, nbar=4):
    """Taylor windows along pulses and samples, float32 (fastsar/api.py:54-56)."""
    from scipy.signal.windows import taylor
    return (taylor(P, nbar=nbar, sll=sll, norm=False).astype(np.float32),
            taylor(K, nbar=nbar, sll=sll, norm=False).astype(np.float32))


# ------------------------------------------------------------------------------------------------ exact BP (FP64)

def range_compress(S, nfft):
    """S [p, K] -> [p, nfft] complex128 profiles: samples centred and zero padded, inverse FFT times nfft, zero
    delay at nfft // 2 (the carrier fmin + (K // 2) df removed). (fastsar/bp.py:55-63)"""
    S = np.asarray(S, np.complex128)
    P, K = S.shape
    h = K // 2
    pad = np.zeros((P, nfft), np.complex128)
    pad[:, :K - h] = S[:, h:]
    pad[:, nfft - h:] = S[:, :h]
    return np.fft.fftshift(np.fft.ifft(pad, axis=1), axes=1) * nfft


def bp_points(S, ant, fmin, df, pts, ref=None, upsample=64, interp='linear', chunk=64):
    """FP64 exact backprojection of S [P, K] (already windowed, if at all) onto points [..., 3]: profiles oversampled
    upsample times (nfft a power of two) in complex128, read linearly or by cubic Lagrange interpolation at the exact
    float64 range of every pixel and pulse. -> complex128 [...]."""
    S = np.asarray(S)
    ant = np.asarray(ant, np.float64)
    P, K = S.shape
    ref = np.linalg.norm(ant, axis=1) if ref is None else np.asarray(ref, np.float64)
    pts = np.asarray(pts, np.float64)
    shape = pts.shape[:-1]
    flat = np.ascontiguousarray(pts.reshape(-1, 3))
    nfft = 1 << int(np.ceil(np.log2(upsample * K)))
    inv_dr = 2.0 * df * nfft / C
    kcyc = 2.0 * (fmin + (K // 2) * df) / C
    ore = np.zeros(len(flat))
    oim = np.zeros(len(flat))
    L = lib()
    for p0 in range(0, P, chunk):
        p1 = min(P, p0 + chunk)
        rc = range_compress(S[p0:p1], nfft)
        L.orc_bp_points(np.ascontiguousarray(rc.real), np.ascontiguousarray(rc.imag), p1 - p0, nfft,
                        np.ascontiguousarray(ant[p0:p1]), np.ascontiguousarray(ref[p0:p1]), flat, len(flat), inv_dr, kcyc,
                        3 if interp == 'cubic' else 1, ore, oim)
    return (ore + 1j * oim).reshape(shape)


def bp_points_weighted(S, ant, fmin, df, pts, weight, upsample=32, interp='cubic'):
    """bp_points with a weight per point and pulse: sum_p weight(pts, [p])[0, n] prof_p(t_pn) exp(2 pi i kcyc dR_pn),
    the image an aperture weight asks for (fastsar/api.py:203-244 approximates it per final subaperture and tile).
    weight(points [n, 3], pulses [m] int) -> [m, n]. -> complex128 [...]."""
    S = np.asarray(S)
    ant = np.asarray(ant, np.float64)
    P, K = S.shape
    ref = np.linalg.norm(ant, axis=1)
    pts = np.asarray(pts, np.float64)
    shape = pts.shape[:-1]
    flat = np.ascontiguousarray(pts.reshape(-1, 3))
    nfft = 1 << int(np.ceil(np.log2(upsample * K)))
    inv_dr = 2.0 * df * nfft / C
    kcyc = 2.0 * (fmin + (K // 2) * df) / C
    out = np.zeros(len(flat), np.complex128)
    L = lib()
    for p0 in range(0, P, 64):
        p1 = min(P, p0 + 64)
        rc = range_compress(S[p0:p1], nfft)
        w = np.asarray(weight(flat, np.arange(p0, p1)), np.float64)
        for i in range(p1 - p0):
            ore, oim = np.zeros(len(flat)), np.zeros(len(flat))
            L.orc_bp_points(np.ascontiguousarray(rc[i:i + 1].real), np.ascontiguousarray(rc[i:i + 1].imag), 1, nfft,
                            np.ascontiguousarray(ant[p0 + i:p0 + i + 1]), np.ascontiguousarray(ref[p0 + i:p0 + i + 1]),
                            flat, len(flat), inv_dr, kcyc, 3 if interp == 'cubic' else 1, ore, oim)
            out += w[i] * (ore + 1j * oim)
    return out.reshape(shape)


def direct_sum(S, ant, fmin, df, pts, ref=None):
    """The interpolation-free image sum_p sum_k S[p, k] exp(+j 4 pi f_k dR / c) (complex128), for tiny cases."""
    S = np.asarray(S, np.complex128)
    ant = np.asarray(ant, np.float64)
    P, K = S.shape
    ref = np.linalg.norm(ant, axis=1) if ref is None else np.asarray(ref, np.float64)
    f = fmin + df * np.arange(K)
    flat = np.asarray(pts, np.float64).reshape(-1, 3)
    out = np.zeros(len(flat), np.complex128)
    for p in range(P):
        dR = np.linalg.norm(flat - ant[p], axis=1) - ref[p]
        out += np.exp(4j * np.pi / C * dR[:, None] * f[None, :]) @ S[p]
    return out.reshape(np.shape(pts)[:-1])


def range_gate(S, g, win=None):
    """S [P, K] gated in range by the plan g (npl, Kp, N, k0, k1, taper) in complex128: windowed by win = (wp, wk)
    when given, padded by npl samples below to Kp, inverse-transformed (numpy's 1 / Kp scale), cut to the first
    N - N/2 and last N/2 range bins, transformed back and kept over [k0, k1) times the taper. -> [P, k1 - k0]
    complex128. (fastsar/api.py:546-563)"""
    S = np.asarray(S, np.complex128)
    P, K = S.shape
    npl, Kp, N, k0, k1 = g['npl'], g['Kp'], g['N'], g['k0'], g['k1']
    h = N // 2
    z = np.zeros((P, Kp), np.complex128)
    z[:, npl:npl + K] = S if win is None else S * (np.asarray(win[0], np.float64)[:, None] * np.asarray(win[1], np.float64)[None, :])
    q = np.fft.ifft(z, axis=1)
    u = np.fft.fft(np.concatenate([q[:, :N - h], q[:, Kp - h:]], axis=1), axis=1)
    return u[:, k0:k1] * np.asarray(g['taper'], np.float64)


def plane_points(nx, ny, spx, spy, e1=(1.0, 0.0, 0.0), e2=(0.0, 1.0, 0.0), center=None):
    """Grid points x_ij = (i - nx/2) spx e1 + (j - ny/2) spy e2 (+ center): [nx, ny, 3]. (fastsar/bp.py:351-361)"""
    e1, e2 = np.asarray(e1, np.float64), np.asarray(e2, np.float64)
    X, Y = np.meshgrid((np.arange(nx) - nx / 2.0) * spx, (np.arange(ny) - ny / 2.0) * spy, indexing='ij')
    p = X[..., None] * e1 + Y[..., None] * e2
    return p if center is None else p + np.asarray(center, np.float64)


# ------------------------------------------------------------------------------------------------ FFBP stages

def rot_fir_k(xre, xim, c0, sl, fr, D, Ko, kc):
    """float64 stage: x [Np, P, K] planes, c0/sl [Np, C, P] -> y planes [Np C, P, Ko]. fr: the filter dict (None: D = 1)."""
    Np, P, K = xre.shape
    Cn = c0.shape[1]
    taps = np.ones(1) if fr is None else np.asarray(fr['kern'], np.float64)
    L, pl = (1, 0) if fr is None else (fr['L'], fr['pl'])
    yre = np.empty((Np * Cn, P, Ko))
    yim = np.empty((Np * Cn, P, Ko))
    lib().orc_rot_fir_k(np.ascontiguousarray(xre, np.float64), np.ascontiguousarray(xim, np.float64), Np, P, K,
                        np.ascontiguousarray(c0, np.float64), np.ascontiguousarray(sl, np.float64), Cn,
                        np.ascontiguousarray(taps), L, pl, D, Ko, float(kc), yre, yim)
    return yre, yim


def fir_p(yre, yim, fr, D, Po):
    B, P, Ko = yre.shape
    zre = np.empty((B, Po, Ko))
    zim = np.empty((B, Po, Ko))
    lib().orc_fir_p(np.ascontiguousarray(yre, np.float64), np.ascontiguousarray(yim, np.float64), B, P, Ko,
                    np.ascontiguousarray(fr['kern'], np.float64), fr['L'], fr['pl'], D, Po, zre, zim)
    return zre, zim


def final_tiles(dre, dim, ux, uy, dlx, dly, a0, a1, ucx, ucy, rc, fc2, T, wts=None):
    B, Pf, Qf = dre.shape
    ore = np.empty((B, T, T))
    oim = np.empty((B, T, T))
    w = [None] * 3 if wts is None else [np.ascontiguousarray(v, np.float64) for v in wts]
    lib().orc_final_tiles(np.ascontiguousarray(dre, np.float64), np.ascontiguousarray(dim, np.float64), B, Pf, Qf, T,
                          np.ascontiguousarray(ux, np.float64), np.ascontiguousarray(uy, np.float64),
                          np.ascontiguousarray(dlx, np.float64), np.ascontiguousarray(dly, np.float64), a0, a1,
                          np.ascontiguousarray(ucx, np.float64), np.ascontiguousarray(ucy, np.float64),
                          np.ascontiguousarray(rc, np.float64), fc2,
                          *[None if v is None else v.ctypes.data for v in w], ore, oim)
    return ore, oim


def ref_rot_fir_k(xre, xim, c0, sl, fr, D, Ko, kc):
    """The reference's float32 stage (its numpy rotation when D = 1, fastsar/ffbp_cpu.py:79-87)."""
    Np, P, K = xre.shape
    Cn = c0.shape[1]
    if fr is None:
        k = np.arange(K, dtype=np.float64) - kc
        cyc = c0[:, :, :, None].astype(np.float64) + k[None, None, None, :] * sl[:, :, :, None]
        ang = (cyc - np.rint(cyc)) * (2 * np.pi)
        cs, sn = np.cos(ang).astype(np.float32), np.sin(ang).astype(np.float32)
        pre, pim = xre.astype(np.float32), xim.astype(np.float32)
        return (np.ascontiguousarray((pre[:, None] * cs - pim[:, None] * sn).reshape(Np * Cn, P, K)),
                np.ascontiguousarray((pre[:, None] * sn + pim[:, None] * cs).reshape(Np * Cn, P, K)))
    yre = np.empty((Np * Cn, P, Ko), np.float32)
    yim = np.empty((Np * Cn, P, Ko), np.float32)
    ref_lib().rot_fir_k(np.ascontiguousarray(xre, np.float32), np.ascontiguousarray(xim, np.float32), Np, P, K,
                        np.ascontiguousarray(c0, np.float32), np.ascontiguousarray(sl, np.float32), Cn,
                        np.ascontiguousarray(fr['kern'], np.float32), fr['L'], fr['pl'], D, Ko, float(kc), yre, yim)
    return yre, yim


def ref_fir_p(yre, yim, fr, D, Po):
    B, P, Ko = yre.shape
    zre = np.empty((B, Po, Ko), np.float32)
    zim = np.empty((B, Po, Ko), np.float32)
    ref_lib().fir_p(np.ascontiguousarray(yre, np.float32), np.ascontiguousarray(yim, np.float32), B, P, Ko,
                    np.ascontiguousarray(fr['kern'], np.float32), fr['L'], fr['pl'], D, Po, zre, zim)
    return zre, zim


def ref_final_tiles(dre, dim, ux, uy, dlx, dly, a0, a1, ucx, ucy, rc, fc2, T, wts=None):
    """The reference's float32 final stage; it supports T = 16 or 32 only on x86 (fastsar/ffbp_cpu.cpp:462-467)."""
    if T not in (16, 32):
        raise ValueError(f"the reference's CPU final stage takes T = 16 or 32, not {T}")
    B, Pf, Qf = dre.shape
    ore = np.empty((B, T, T), np.float32)
    oim = np.empty((B, T, T), np.float32)
    w = [None] * 3 if wts is None else [np.ascontiguousarray(v, np.float32) for v in wts]
    ref_lib().final_tiles(np.ascontiguousarray(dre, np.float32), np.ascontiguousarray(dim, np.float32), B, Pf, Qf, T,
                          np.ascontiguousarray(ux, np.float64), np.ascontiguousarray(uy, np.float64),
                          np.ascontiguousarray(dlx, np.float64), np.ascontiguousarray(dly, np.float64), a0, a1,
                          np.ascontiguousarray(ucx, np.float64), np.ascontiguousarray(ucy, np.float64),
                          np.ascontiguousarray(rc, np.float64), fc2, ore, oim,
                          *[None if v is None else v.ctypes.data for v in w])
    return ore, oim


def final_geometry(fu, fr0, cen, e1, e2):
    """Per tile b and final pulse p: the unit vector from the tile centre to the antenna along e1, e2 (ux, uy [B, Pf])
    and per tile the normalized aperture-mean direction (ucx, ucy) and mean range rc. (fastsar/ffbp_cpu.py:187-199)"""
    e1, e2 = np.asarray(e1, np.float64), np.asarray(e2, np.float64)
    en = np.cross(e1, e2)
    w = fr0[None, :, None] * fu[None, :, :] - cen[:, None, :]
    wn = np.sqrt((w * w).sum(2))
    ux, uy, uz = (w @ e1) / wn, (w @ e2) / wn, (w @ en) / wn
    mx, my, mz = ux.mean(1), uy.mean(1), uz.mean(1)
    mn = np.sqrt(mx * mx + my * my + mz * mz)
    return np.ascontiguousarray(ux), np.ascontiguousarray(uy), mx / mn, my / mn, wn.mean(1)


def ffbp_image(plan, coll, S, stages='ref', device_phases=None):
    """The CPU former's image of the (windowed) history S [P, K] for a plan and its collection arrays: the level-0
    children of the history one at a time, each carried through the remaining levels and the final stage, the
    tiles assembled and cropped. stages='ref': the reference's float32 kernels; 'oracle': the float64 ones.
    device_phases(u, r0, d, refs, k0c, k1): the planner's float64 phases of levels >= 2.
    (fastsar/ffbp_cpu.py:158-291)"""
    from fastsar_amd import plan as planner
    device_phases = device_phases or planner.device_phases
    r64 = stages == 'oracle'
    rfk, rfp, rfin = (rot_fir_k, fir_p, final_tiles) if r64 else (ref_rot_fir_k, ref_fir_p, ref_final_tiles)
    levels, T, nlev = plan['levels'], plan['T'], len(plan['levels'])
    fin = plan['final']
    a0, a1, fc2 = 2.0 * fin['f0'] / C, 2.0 * fin['df'] / C, 2.0 * fin['fc'] / C
    dlx = (np.arange(T) - (T - 1) / 2.0) * plan['spx']
    dly = (np.arange(T) - (T - 1) / 2.0) * plan['spy']
    fu, fr0 = coll['final']['u'], coll['final']['r0']
    fcen = fin['cen']
    sx0, sy0 = levels[0]['sx'], levels[0]['sy']
    G = sx0 * sy0
    mx, my = plan['Nx'] // sx0, plan['Ny'] // sy0
    shape_g = [s for lv in levels[1:] for s in (lv['sx'], lv['sy'])] + [T, T]
    perm_g = [2 * i for i in range(nlev - 1)] + [2 * (nlev - 1)] + [2 * i + 1 for i in range(nlev - 1)] + [2 * (nlev - 1) + 1]

    def children(a, b, c0, sl, lv):
        y = rfk(a, b, c0, sl, lv['fir_k'], lv['Dk'], lv['Ko'], (lv['K'] - 1) / 2.0)
        if lv['Dp'] > 1:
            y = rfp(y[0], y[1], lv['fir_p'], lv['Dp'], lv['Po'])
        return y

    S = np.asarray(S)
    dt = np.float64 if r64 else np.float32
    xre, xim = np.ascontiguousarray(S.real, dt)[None], np.ascontiguousarray(S.imag, dt)[None]
    full = np.empty((plan['Nx'], plan['Ny']), np.complex128 if r64 else np.complex64)
    c0all, slall = coll['levels'][0]['c0'], coll['levels'][0]['slope']
    for g in range(G):
        a, b = children(xre, xim, c0all[0, g][None, None], slall[0, g][None, None], levels[0])
        for i in range(1, nlev):
            lv, cl = levels[i], coll['levels'][i]
            if i < planner.HOST_LEVELS:
                c0, sl = cl['c0'][g][None], cl['slope'][g][None]
            else:
                refs = lv['ref'].reshape(G, a.shape[0], 3)[g]
                k0c, k1 = planner.level_phase_constants(lv)
                c0, sl = device_phases(cl['u'], cl['r0'], lv['d'], refs, k0c, k1)
            a, b = children(a[:, None] if a.ndim == 2 else a, b[:, None] if b.ndim == 2 else b, c0, sl, lv)
        cen = fcen.reshape(G, -1, 3)[g]
        ux, uy, ucx, ucy, rc = final_geometry(fu, fr0, cen, plan['e1'], plan['e2'])
        re, im = rfin(a, b, ux, uy, dlx, dly, a0, a1, ucx, ucy, rc, fc2, T)
        blk = re.reshape(shape_g).transpose(perm_g).reshape(mx, my) + 1j * im.reshape(shape_g).transpose(perm_g).reshape(mx, my)
        x, y = g // sy0, g % sy0
        full[x * mx:(x + 1) * mx, y * my:(y + 1) * my] = blk
    ox, oy, nx, ny = plan['ox'], plan['oy'], plan['nx'], plan['ny']
    return full[ox:ox + nx, oy:oy + ny]


class RefCpuFormer:
    """The reference's CPU former (fastsar/ffbp_cpu.py:158-291) over its compiled kernels: level-0 children in
    groups of up to 8, the windowed complex64 history read in place (rot_fir_k_cplx), each child carried through the
    remaining levels (levels >= 2 with float64 NumPy phases) and final_tiles, the blocks written into the image.
    form(S, children=...) forms only the listed level-0 children (the CPU baseline's bounded sample)."""

    def __init__(self, plan, coll):
        from fastsar_amd import plan as planner
        self.plan, self.coll, self.planner = plan, coll, planner
        if plan['T'] not in (16, 32):
            raise ValueError("the reference's CPU former takes T = 16 or 32")
        lv = plan['levels']
        self.G = lv[0]['C']

    def window(self, S, wp, wk):
        """One complex64 copy windowed in row blocks (fastsar/api.py:512-519)."""
        out = np.empty(S.shape, np.complex64)
        for p0 in range(0, S.shape[0], 4096):
            sl = slice(p0, min(S.shape[0], p0 + 4096))
            np.multiply(S[sl], (wp[sl, None] * wk[None, :]).astype(np.float32), out=out[sl])
        return out

    def __call__(self, S, children=None, ng=8):
        plan, coll, planner = self.plan, self.coll, self.planner
        L = ref_lib()
        levels, T, nlev = plan['levels'], plan['T'], len(plan['levels'])
        fin = plan['final']
        a0, a1, fc2 = 2.0 * fin['f0'] / C, 2.0 * fin['df'] / C, 2.0 * fin['fc'] / C
        dlx = np.ascontiguousarray((np.arange(T) - (T - 1) / 2.0) * plan['spx'])
        dly = np.ascontiguousarray((np.arange(T) - (T - 1) / 2.0) * plan['spy'])
        fu, fr0, fcen = coll['final']['u'], coll['final']['r0'], fin['cen']
        sx0, sy0 = levels[0]['sx'], levels[0]['sy']
        G = self.G
        mx, my = plan['Nx'] // sx0, plan['Ny'] // sy0
        shape_g = [s for lv in levels[1:] for s in (lv['sx'], lv['sy'])] + [T, T]
        perm_g = [2 * i for i in range(nlev - 1)] + [2 * (nlev - 1)] + [2 * i + 1 for i in range(nlev - 1)] + [2 * (nlev - 1) + 1]
        S = np.ascontiguousarray(S, np.complex64)
        lv0 = levels[0]
        if lv0['Dk'] <= 1:
            raise ValueError('the in-place complex64 first level needs range decimation at level 0')
        full = np.empty((plan['Nx'], plan['Ny']), np.complex64)
        todo = list(range(G)) if children is None else list(children)

        def kids(a, b, c0, sl, lv):
            y = ref_rot_fir_k(a, b, c0, sl, lv['fir_k'], lv['Dk'], lv['Ko'], (lv['K'] - 1) / 2.0)
            return ref_fir_p(y[0], y[1], lv['fir_p'], lv['Dp'], lv['Po']) if lv['Dp'] > 1 else y

        for g0 in range(0, len(todo), ng):
            gs = todo[g0:g0 + ng]
            c0 = np.ascontiguousarray(coll['levels'][0]['c0'][0, gs][None])
            sl = np.ascontiguousarray(coll['levels'][0]['slope'][0, gs][None])
            fk = lv0['fir_k']
            yre = np.empty((len(gs), lv0['P'], lv0['Ko']), np.float32)
            yim = np.empty((len(gs), lv0['P'], lv0['Ko']), np.float32)
            L.rot_fir_k_cplx(S.view(np.float32).reshape(S.shape[0], -1), 1, lv0['P'], lv0['K'], c0, sl, len(gs),
                             np.ascontiguousarray(fk['kern'], np.float32), fk['L'], fk['pl'], lv0['Dk'], lv0['Ko'],
                             (lv0['K'] - 1) / 2.0, yre, yim)
            A, B = (ref_fir_p(yre, yim, lv0['fir_p'], lv0['Dp'], lv0['Po']) if lv0['Dp'] > 1 else (yre, yim))
            for k, g in enumerate(gs):
                a, b = A[k][None], B[k][None]
                for i in range(1, nlev):
                    lv, cl = levels[i], coll['levels'][i]
                    if i < planner.HOST_LEVELS:
                        c0, sl = cl['c0'][g][None], cl['slope'][g][None]
                    else:
                        refs = lv['ref'].reshape(G, a.shape[0], 3)[g]
                        k0c, k1 = planner.level_phase_constants(lv)
                        c0, sl = planner.device_phases(cl['u'], cl['r0'], lv['d'], refs, k0c, k1)
                    a, b = kids(a, b, c0, sl, lv)
                cen = fcen.reshape(G, -1, 3)[g]
                ux, uy, ucx, ucy, rc = final_geometry(fu, fr0, cen, plan['e1'], plan['e2'])
                re, im = ref_final_tiles(a, b, ux, uy, dlx, dly, a0, a1, ucx, ucy, rc, fc2, T)
                blk = (re.reshape(shape_g).transpose(perm_g).reshape(mx, my)
                       + 1j * im.reshape(shape_g).transpose(perm_g).reshape(mx, my))
                x, y = g // sy0, g % sy0
                full[x * mx:(x + 1) * mx, y * my:(y + 1) * my] = blk
        ox, oy, nx, ny = plan['ox'], plan['oy'], plan['nx'], plan['ny']
        return full[ox:ox + nx, oy:oy + ny]


def exact_tiles(win, lo, ant, ref, cen, ntx, nty, tx, ty, nx, ny, spx, spy, e1, e2, inv_dr, kcyc, cubic):
    """float64 ExactFormer tiles on cropped profiles win [P, W] complex -> image [nx, ny] complex128."""
    P, W = win.shape
    ore = np.zeros(nx * ny)
    oim = np.zeros(nx * ny)
    lib().orc_exact_tiles(np.ascontiguousarray(win.real, np.float64), np.ascontiguousarray(win.imag, np.float64), P, W,
                          np.ascontiguousarray(lo, np.int32), np.ascontiguousarray(ant, np.float64),
                          np.ascontiguousarray(ref, np.float64), np.ascontiguousarray(cen, np.float64), ntx, nty, tx, ty,
                          nx, ny, spx, spy, np.asarray(e1, np.float64), np.asarray(e2, np.float64), inv_dr, kcyc,
                          int(cubic), ore, oim)
    return (ore + 1j * oim).reshape(nx, ny)


# ------------------------------------------------------------------------------------------------ tap network

HID = 64


def _tapnet_split(params, T):
    p = np.asarray(params, np.float64)
    o = [0]

    def take(n, shape):
        a = p[o[0]:o[0] + n].reshape(shape)
        o[0] += n
        return a
    W1, b1 = take(2 * HID, (HID, 2)), take(HID, (HID,))
    W2, b2 = take(HID * HID, (HID, HID)), take(HID, (HID,))
    W3, b3 = take(T * HID, (T, HID)), take(T, (T,))
    return W1, b1, W2, b2, W3, b3


def tapnet_forward(params, T, pts):
    """The tap network in float64: pts [n, 2] -> (out [n, T], h1, h2)."""
    W1, b1, W2, b2, W3, b3 = _tapnet_split(params, T)
    h1 = np.tanh(np.asarray(pts, np.float64) @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3.T + b3, h1, h2


def tapnet_backward(params, T, pts, dout):
    """d loss / d params [nparam] (float64, the device's parameter order) from d loss / d out [n, T], by the chain
    rule through tanh (1 - h^2)."""
    W1, b1, W2, b2, W3, b3 = _tapnet_split(params, T)
    x = np.asarray(pts, np.float64)
    _, h1, h2 = tapnet_forward(params, T, pts)
    g = np.asarray(dout, np.float64)
    gW3, gb3 = g.T @ h2, g.sum(0)
    d2 = (g @ W3) * (1 - h2 ** 2)
    gW2, gb2 = d2.T @ h1, d2.sum(0)
    d1 = (d2 @ W2) * (1 - h1 ** 2)
    gW1, gb1 = d1.T @ x, d1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def ls_loss_grad(out, T, stages, wt=None):
    """The least-squares loss of filters built from network outputs: stages [(D, L, off, R [L, L], v [L], E,
    base [L])]; taps h[D u + ph] = base + out[off + ph, u]; loss_s = (h R h - 2 h v + E) / E; returns (loss [S],
    d (sum wt_s loss_s) / d out)."""
    out = np.asarray(out, np.float64)
    dout = np.zeros_like(out)
    loss = []
    wt = np.ones(len(stages)) if wt is None else np.asarray(wt, np.float64)
    for s, (D, L, off, R, v, E, base) in enumerate(stages):
        r = np.arange(L)
        h = np.asarray(base, np.float64)[:L] + out[off + r % D, r // D]
        loss.append((h @ R @ h - 2 * h @ v + E) / E)
        dout[off + r % D, r // D] += wt[s] * 2 * (R @ h - v) / E
    return np.array(loss), dout
