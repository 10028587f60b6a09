"""The oracle pinned before it is trusted: its float64 stages against the reference's own compiled CPU kernels
(oracle/_ref, built from fastsar/ffbp_cpu.cpp) on identical inputs, against brute-force sums on tiny cases, and
against the mathematical properties the reference's tests rely on (tests/test_bp.py, tests/test_accuracy.py,
tests/test_exact.py). CPU only."""
import numpy as np
import pytest

import oracle
from fastsar_amd import plan as pl, sim

C = pl.C
needs_ref = pytest.mark.skipif(not oracle.have_ref(), reason="the reference's CPU kernels are not built (oracle/_ref)")


def db(a, b):
    return oracle.plain_db(a, b)


# ------------------------------------------------------------------------------------------------ stages vs reference

@needs_ref
@pytest.mark.parametrize('D,pf,Cn,Np,K', [(3, 0.6, 5, 2, 301), (7, 0.5, 16, 1, 777), (2, 0.9, 3, 2, 140)])
def test_rot_fir_k_vs_reference(D, pf, Cn, Np, K):
    rng = np.random.default_rng(D + Cn)
    P = 19
    x = (rng.standard_normal((Np, P, K)) + 1j * rng.standard_normal((Np, P, K))).astype(np.complex64)
    c0 = rng.uniform(-0.5, 0.5, (Np, Cn, P)).astype(np.float32)
    sl = rng.uniform(-0.7, 0.7, (Np, Cn, P)).astype(np.float32)
    fr, _ = pl.decimator_fir(K, D, pf)
    kc = (K - 1) / 2.0
    a = oracle.rot_fir_k(x.real, x.imag, c0, sl, fr, D, fr['n_out'], kc)
    b = oracle.ref_rot_fir_k(x.real, x.imag, c0, sl, fr, D, fr['n_out'], kc)
    e = db(b[0] + 1j * b[1], a[0] + 1j * a[1])
    print(f'rot_fir_k D={D} C={Cn}: reference float32 vs oracle float64 {e:.1f} dB')
    assert e < -95, e


@needs_ref
@pytest.mark.parametrize('D,pf', [(2, 0.5), (11, 0.7), (6, 0.3)])
def test_fir_p_vs_reference(D, pf):
    rng = np.random.default_rng(D)
    B, P, Ko = 3, 500, 41
    fr, _ = pl.decimator_fir(P, D, pf)
    y = (rng.standard_normal((B, P, Ko)) + 1j * rng.standard_normal((B, P, Ko))).astype(np.complex64)
    a = oracle.fir_p(y.real, y.imag, fr, D, fr['n_out'])
    b = oracle.ref_fir_p(y.real, y.imag, fr, D, fr['n_out'])
    A, Bv = a[0] + 1j * a[1], b[0] + 1j * b[1]
    # every output but the first agrees to float32 rounding
    assert db(Bv[:, 1:], A[:, 1:]) < -125
    # output 0: fastsar/ffbp_cpu.cpp:274 computes ilo = (p + pl - L) / D + 1 with C's truncating division, so rows
    # p = r - pl >= 0 with taps r in [L - D + 1, L - 1] are not added to it; the decimation matrix (fastsar/ffbp.py:57-78)
    # and the oracle include them. The difference is exactly their sum.
    r = np.arange(fr['L'] - D + 1, fr['L'])
    p = r - fr['pl']
    keep = p >= 0
    skipped = (fr['kern'][r[keep]][None, :, None] * y[:, p[keep], :]).sum(1)
    assert np.abs((A[:, 0] - Bv[:, 0]) - skipped).max() < 1e-5 * np.abs(A).max()


def _final_case(rng, B, Pf, Qf, T):
    d = (rng.standard_normal((B, Pf, Qf)) + 1j * rng.standard_normal((B, Pf, Qf))).astype(np.complex64)
    ux = rng.uniform(-0.1, 0.1, (B, Pf)) + 0.05
    uy = rng.uniform(0.8, 0.9, (B, Pf))
    geo = np.stack([rng.uniform(0, .1, B), rng.uniform(.8, .9, B), rng.uniform(9e3, 1.1e4, B)], 1)
    f0, df, sp = 9.45e9, 3e6, 0.5
    a0, a1, fc2 = 2 * f0 / C, 2 * df / C, 2 * (f0 + (Qf - 1) / 2 * df) / C
    dl = (np.arange(T) - (T - 1) / 2.0) * sp
    return d, ux, uy, geo, dl, a0, a1, fc2


@needs_ref
@pytest.mark.parametrize('T', [16, 32])
@pytest.mark.parametrize('weights', [False, True])
def test_final_tiles_vs_reference(T, weights):
    rng = np.random.default_rng(T)
    B, Pf, Qf = 3, 17, 29
    d, ux, uy, geo, dl, a0, a1, fc2 = _final_case(rng, B, Pf, Qf, T)
    wts = None
    if weights:
        wts = tuple(v.astype(np.float32) for v in (rng.uniform(0.5, 1.5, (B, Pf)), rng.uniform(-.01, .01, (B, Pf)),
                                                     rng.uniform(-.01, .01, (B, Pf))))
    a = oracle.final_tiles(d.real, d.imag, ux, uy, dl, dl, a0, a1, geo[:, 0], geo[:, 1], geo[:, 2], fc2, T, wts)
    b = oracle.ref_final_tiles(d.real, d.imag, ux, uy, dl, dl, a0, a1, geo[:, 0], geo[:, 1], geo[:, 2], fc2, T, wts)
    e = db(b[0] + 1j * b[1], a[0] + 1j * a[1])
    print(f'final_tiles T={T} weights={weights}: reference vs oracle {e:.1f} dB')
    assert e < -100, e


def test_final_tiles_brute_force():
    """the separable product against the plain double sum over (p, q) of every pixel, tiny tile"""
    rng = np.random.default_rng(3)
    B, Pf, Qf, T = 2, 3, 5, 4
    d, ux, uy, geo, dl, a0, a1, fc2 = _final_case(rng, B, Pf, Qf, T)
    re, im = oracle.final_tiles(d.real, d.imag, ux, uy, dl, dl, a0, a1, geo[:, 0], geo[:, 1], geo[:, 2], fc2, T)
    want = np.zeros((B, T, T), complex)
    for b in range(B):
        for i in range(T):
            for j in range(T):
                s = 0j
                for p in range(Pf):
                    for q in range(Qf):
                        s += d[b, p, q] * np.exp(-2j * np.pi * (ux[b, p] * dl[i] + uy[b, p] * dl[j]) * (a0 + a1 * q))
                los = geo[b, 0] * dl[i] + geo[b, 1] * dl[j]
                qc = (dl[i] ** 2 + dl[j] ** 2 - los ** 2) * fc2 / (2 * geo[b, 2])
                want[b, i, j] = s * np.exp(2j * np.pi * qc)
    assert db(re + 1j * im, want) < -200


def test_rot_fir_k_brute_force():
    """rotation and decimation against the literal double sum, tiny case (a sign or index slip fails it)"""
    rng = np.random.default_rng(5)
    P, K, D, Cn = 2, 40, 3, 2
    x = rng.standard_normal((1, P, K)) + 1j * rng.standard_normal((1, P, K))
    c0 = rng.uniform(-.5, .5, (1, Cn, P))
    sl = rng.uniform(-.3, .3, (1, Cn, P))
    fr, _ = pl.decimator_fir(K, D, 0.5)
    kc = (K - 1) / 2
    yre, yim = oracle.rot_fir_k(x.real, x.imag, c0, sl, fr, D, fr['n_out'], kc)
    want = np.zeros((Cn, P, fr['n_out']), complex)
    for c in range(Cn):
        for p in range(P):
            z = x[0, p] * np.exp(2j * np.pi * (c0[0, c, p] + (np.arange(K) - kc) * sl[0, c, p]))
            F, _ = pl.decimator_matrix(K, D, 0.5)
            want[c, p] = z @ F
    assert db(yre + 1j * yim, want) < -200


# ------------------------------------------------------------------------------------------------ FP64 backprojection

def _scene(r0=5e3, n_t=6, seed=3, res=1.0, scene=30.0):
    rng = np.random.default_rng(seed)
    col = sim.make_collect(res=res, scene=scene, r0=r0)
    tg = np.stack([rng.uniform(-10, 10, n_t), rng.uniform(-10, 10, n_t), np.zeros(n_t)], 1)
    amp = rng.standard_normal(n_t) + 1j * rng.standard_normal(n_t)
    return col, tg, amp, sim.simulate_brute(col, tg, amp)


def test_bp_points_vs_direct_sum():
    """oversampled profiles read by interpolation against the interpolation-free sum over pulses and samples"""
    col, tg, amp, S = _scene()
    pts = np.concatenate([tg, tg + [0.3, -0.2, 0.0], np.array([[1.1, 2.2, 0.0], [-4.0, 3.0, 0.5]])])
    direct = oracle.direct_sum(S, col.ant, col.fmin, col.df, pts)
    lin = oracle.bp_points(S, col.ant, col.fmin, col.df, pts, upsample=64, interp='linear')
    cub = oracle.bp_points(S, col.ant, col.fmin, col.df, pts, upsample=64, interp='cubic')
    el, ec = db(lin, direct), db(cub, direct)
    print(f'bp_points vs direct sum: linear 64x {el:.1f} dB, cubic 64x {ec:.1f} dB')
    assert el < -80 and ec < -140


def test_bp_points_vs_numpy_reference():
    """the C kernel against a complex128 NumPy backprojection of the same profiles (tests/test_bp.py:23-35)"""
    col, tg, amp, S = _scene(r0=600e3)
    pts = np.stack([np.random.default_rng(1).uniform(-15, 15, 40), np.random.default_rng(2).uniform(-15, 15, 40),
                    np.zeros(40)], 1)
    P, K = S.shape
    nfft = 1 << int(np.ceil(np.log2(16 * K)))
    rc = oracle.range_compress(S, nfft)
    want = np.zeros(len(pts), complex)
    for p in range(P):
        dR = np.linalg.norm(pts - col.ant[p], axis=1) - np.linalg.norm(col.ant[p])
        t = dR * 2 * col.df * nfft / C + nfft // 2
        i = np.floor(t).astype(int)
        w = t - i
        want += (rc[p, i % nfft] * (1 - w) + rc[p, (i + 1) % nfft] * w) * np.exp(2j * np.pi * 2 * (col.fmin + (K // 2) * col.df) / C * dR)
    got = oracle.bp_points(S, col.ant, col.fmin, col.df, pts, upsample=16)
    assert db(got, want) < -250


def test_bp_points_weighted():
    """the per-point, per-pulse weighted backprojection: unit weights give bp_points; weights that depend on the
    pulse alone give bp_points of the history times those weights; the image is linear in the weight; a weight that
    is one at a single point and zero elsewhere leaves only that point's value"""
    col, tg, amp, S = _scene()
    P = S.shape[0]
    pts = np.concatenate([tg, np.array([[1.1, 2.2, 0.0], [-4.0, 3.0, 0.5], [0.3, -7.0, 0.0]])])
    plain = oracle.bp_points(S, col.ant, col.fmin, col.df, pts, upsample=16, interp='cubic')
    ones = oracle.bp_points_weighted(S, col.ant, col.fmin, col.df, pts, lambda q, p: np.ones((len(p), len(q))), upsample=16)
    assert db(ones, plain) < -250
    hw = 0.5 - 0.5 * np.cos(2 * np.pi * (np.arange(P) + 0.5) / P)
    pw = oracle.bp_points_weighted(S, col.ant, col.fmin, col.df, pts, lambda q, p: np.repeat(hw[p][:, None], len(q), 1),
                                   upsample=16)
    assert db(pw, oracle.bp_points(S * hw[:, None], col.ant, col.fmin, col.df, pts, upsample=16, interp='cubic')) < -250

    def wa(q, p):
        return np.outer(np.cos(np.asarray(p) / 7.0), 1.0 + 0.01 * q[:, 0])

    def wb(q, p):
        return np.outer(np.asarray(p) / P, np.exp(-0.02 * q[:, 1]))
    both = oracle.bp_points_weighted(S, col.ant, col.fmin, col.df, pts, lambda q, p: wa(q, p) + wb(q, p), upsample=16)
    sep = (oracle.bp_points_weighted(S, col.ant, col.fmin, col.df, pts, wa, upsample=16) +
           oracle.bp_points_weighted(S, col.ant, col.fmin, col.df, pts, wb, upsample=16))
    assert db(both, sep) < -250
    one = oracle.bp_points_weighted(S, col.ant, col.fmin, col.df, pts,
                                    lambda q, p: np.repeat((np.arange(len(q)) == 2)[None, :].astype(float), len(p), 0),
                                    upsample=16)
    assert np.all(one[np.arange(len(pts)) != 2] == 0) and abs(one[2] / plain[2] - 1) < 1e-12


def test_bp_oversampling_rate():
    """linear reads: error against a 128x reference falls about 12 dB per doubling (fastsar/bp.py:15-16)"""
    col, tg, amp, S = _scene(r0=4e3)
    pts = oracle.plane_points(24, 24, 0.5, 0.5, (0.0, 1.0, 0.0), (1.0, 0.0, 0.0))
    ref = oracle.bp_points(S, col.ant, col.fmin, col.df, pts, upsample=128)
    errs = [oracle.rel_db(oracle.bp_points(S, col.ant, col.fmin, col.df, pts, upsample=u), ref) for u in (4, 8, 16)]
    print('oversampling 4, 8, 16x:', [f'{e:.1f}' for e in errs])
    assert errs[1] < errs[0] - 10 and errs[2] < errs[1] - 10


def test_range_periodic():
    """a scatterer 0.7 of the unambiguous range from the reference focuses as one at the reference
    (tests/test_bp.py:142-154)"""
    col = sim.make_collect(res=1.0, scene=30.0, r0=5e3)
    L = C / (2 * col.df)
    vals = []
    for off in (0.0, 0.7 * L):
        x = np.array([[off * np.cos(np.deg2rad(30)), 0.0, 0.0]])
        S = sim.simulate_brute(col, x, np.ones(1))
        vals.append(abs(oracle.bp_points(S, col.ant, col.fmin, col.df, x)[0]))
    assert abs(vals[1] / vals[0] - 1) < 0.02


def test_gain():
    """a unit scatterer at a pixel sums to P K (the image is the plain sum over pulses and samples)"""
    col, tg, amp, S = _scene(n_t=1)
    v = oracle.bp_points(S / amp[0], col.ant, col.fmin, col.df, tg)[0]
    assert abs(abs(v) / (S.shape[0] * S.shape[1]) - 1) < 0.01


@pytest.mark.parametrize('r0', [20e3, 600e3])
def test_exact_tiles_vs_bp(r0, monkeypatch):
    """ExactFormer's tiled algorithm (float64, on the hip former's own host plan: tile, centres, reachable bins)
    against exact backprojection: cubic < -65 dB (tests/test_exact.py:40-50)"""
    from fastsar_amd.exact import ExactFormer
    monkeypatch.setattr('fastsar_amd.api._backend', lambda b: 'hip')     # plan on the host; no device is touched
    rng = np.random.default_rng(5)
    col = sim.make_collect(res=0.5, scene=60.0, r0=r0)
    tg = np.stack([rng.uniform(-25, 25, 30), rng.uniform(-25, 25, 30), np.zeros(30)], 1)
    S = sim.simulate_brute(col, tg, rng.standard_normal(30) + 1j * rng.standard_normal(30)).astype(np.complex64)
    grid = dict(nx=100, ny=120, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    f = ExactFormer(col.ant, col.fmin, col.df, S.shape[1], **grid)
    wp, wk = oracle.window(*S.shape)
    Sw = S * wp[:, None] * wk[None, :]
    rc = oracle.range_compress(Sw, f.nfft)
    idx = (f.blo[:, None] + np.arange(f.W)[None, :]) % f.nfft
    win = np.take_along_axis(rc, idx, axis=1)
    img = oracle.exact_tiles(win, f.lo, col.ant, np.linalg.norm(col.ant, axis=1), f.cen, f.ntx, f.nty, f.tx, f.ty,
                             100, 120, 0.5, 0.5, grid['e1'], grid['e2'], f.inv_dr, f.kcyc, True)
    ref = oracle.bp_points(Sw, col.ant, col.fmin, col.df, oracle.plane_points(**grid), upsample=64)
    e = oracle.rel_db(img, ref)
    print(f'exact tiles at {r0 / 1e3:.0f} km (tile {f.tile}): {e:.1f} dB')
    assert e < -65


# ------------------------------------------------------------------------------------------------ whole FFBP images

@pytest.fixture(scope='module')
def small_scene():
    rng = np.random.default_rng(1)
    col = sim.make_collect(res=0.5, scene=60.0, r0=5e3)
    pos = np.stack([rng.uniform(-25, 25, 40), rng.uniform(-25, 25, 40), np.zeros(40)], 1)
    amp = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    return col, sim.simulate_brute(col, pos, amp).astype(np.complex64)


@needs_ref
@pytest.mark.parametrize('nlev', [2, 3])
def test_ffbp_reference_vs_oracle(small_scene, nlev):
    """the reference's CPU former (its kernels, this orchestration) against the float64 stages (< -85 dB; the
    reference's own CPU vs dense check is < -70 dB, tests/test_ffbp_cpu.py:38)"""
    col, S = small_scene
    p = pl.make_plan(col, 128, 128, 0.5, 0.5, T=16, nlev=nlev, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    coll = pl.collection_arrays(p, col.ant)
    a = oracle.ffbp_image(p, coll, S, 'ref')
    b = oracle.ffbp_image(p, coll, S, 'oracle')
    e = db(a, b)
    print(f'nlev {nlev}: reference CPU former vs float64 oracle {e:.1f} dB')
    assert e < -85 and abs(np.abs(a).max() / np.abs(b).max() - 1) < 1e-3


@needs_ref
def test_ref_cpu_former_groups(small_scene):
    """RefCpuFormer (groups of first-level children, complex64 read in place) equals the one-child-at-a-time form
    bit for bit, and its children subsets assemble to the whole (tests/test_units.py:165-185)"""
    col, S = small_scene
    p = pl.make_plan(col, 128, 128, 0.5, 0.5, T=16, nlev=2, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    coll = pl.collection_arrays(p, col.ant)
    f = oracle.RefCpuFormer(p, coll)
    a = f(S, ng=8)
    b = f(S, ng=1)
    assert np.array_equal(a, b)
    assert db(a, oracle.ffbp_image(p, coll, S, 'ref')) < -130


@pytest.mark.parametrize('r0', [1e3, 4e3, 16e3])
def test_final_stage_prediction(r0):
    """T=16 and T=32 images against exact backprojection beside final_phase_error's prediction: within 2 dB where
    the prediction is above -55 dB, T=16 not worse than T=32 (tests/test_accuracy.py:29-54)"""
    n, sp = 128, 0.5
    e1, e2 = (0.0, 1.0, 0.0), (1.0, 0.0, 0.0)
    rng = np.random.default_rng(7)
    col = sim.make_collect(res=0.5, scene=n * sp, r0=r0)
    tg = np.stack([rng.uniform(-25, 25, 30), rng.uniform(-25, 25, 30), np.zeros(30)], 1)
    S = sim.simulate_brute(col, tg, rng.standard_normal(30) + 1j * rng.standard_normal(30)).astype(np.complex64)
    wp, wk = oracle.window(*S.shape)
    Sw = (S * wp[:, None] * wk[None, :]).astype(np.complex64)
    ref = oracle.bp_points(Sw, col.ant, col.fmin, col.df, oracle.plane_points(n, n, sp, sp, e1, e2), upsample=16)
    row = []
    for T in (16, 32):
        p = pl.make_plan(col, n, n, sp, sp, T=T, nlev=3, e1=e1, e2=e2)
        img = oracle.ffbp_image(p, pl.collection_arrays(p, col.ant), Sw, 'oracle')
        pred = pl.final_phase_error(col.ant, col.fmin + col.K * col.df, n, n, sp, sp, e1, e2, T)
        row.append((T, pred, oracle.rel_db(img, ref)))
    print(f'{r0 / 1e3:.0f} km: ' + '; '.join(f'T={T} measured {m:.1f} predicted {p_:.1f}' for T, p_, m in row))
    for T, pred, meas in row:
        assert not (pred > -55 and abs(meas - pred) > 2.0), (T, pred, meas)
        assert meas < max(pred + 2.0, -55.0)
    if row[1][1] > -55:
        assert row[0][2] <= row[1][2]


def test_window_is_scipy_taylor():
    from scipy.signal.windows import taylor
    wp, wk = oracle.window(100, 64)
    assert np.array_equal(wp, taylor(100, nbar=4, sll=35, norm=False).astype(np.float32))
    assert np.array_equal(wk, taylor(64, nbar=4, sll=35, norm=False).astype(np.float32))


# ------------------------------------------------------------------------------------------------ range gate

def wide_band(K=6000, P=96, ncl=60, seed=3):
    """An image narrow in range from a wide-band collection: 0.5 m resolution at 30 km, 40 degrees grazing, points
    in a 19 m image and clutter over 2 km of range around it (the reference's wide_band fixture,
    tests/test_api.py:132-148, at a quarter of its samples and pulses)."""
    rng = np.random.default_rng(seed)
    res, r0 = 0.5, 30e3
    df = C / (2 * res) / K
    y = (np.arange(P) - (P - 1) / 2) * (r0 * C / 9.6e9 / (2 * res) / P)
    psi = np.radians(40.0)
    ant = np.stack([np.full(P, -r0 * np.cos(psi)), y, np.full(P, r0 * np.sin(psi))], 1)
    col = sim.Collect(fmin=9.6e9 - (K // 2) * df, df=df, K=K, ant=ant, res=res)
    pts = np.array([[0.0, 0.0, 0], [5.3, -3.1, 0], [-7.2, 6.4, 0]])
    cl = np.stack([(rng.random(ncl) - 0.5) * 2000, (rng.random(ncl) - 0.5) * 30, np.zeros(ncl)], 1)
    amp = np.concatenate([np.full(3, 30.0), (rng.standard_normal(ncl) + 1j * rng.standard_normal(ncl)) / np.sqrt(2)])
    S = sim.simulate_brute(col, np.concatenate([pts, cl]), amp).astype(np.complex64)
    grid = dict(nx=48, ny=48, spx=0.4, spy=0.4, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    return col, S, grid


@pytest.fixture(scope='module')
def gated_scene():
    from fastsar_amd import gate
    col, S, grid = wide_band()
    return col, S, grid, gate.gate_plan(col.ant, None, col.fmin, col.df, col.K, **grid)


def _gate_images(col, S, grid, g, Sg, win):
    pts = oracle.plane_points(**grid)
    Sw = S if win is None else S * (win[0][:, None] * win[1][None, :])
    full = oracle.bp_points(Sw, col.ant, col.fmin, col.df, pts, upsample=16, interp='cubic')
    gated = oracle.bp_points(Sg, col.ant, g['fmin'], g['df'], pts, upsample=16, interp='cubic')
    return gated, full


@pytest.mark.parametrize('window', [False, True])
def test_range_gate_image(gated_scene, window):
    """The gated history forms the image of the whole one: float64 exact backprojection of both onto the image's
    grid, compared with no fitted gain, so that the range cut, the direction of either transform and the Kp / N
    scale are all pinned (the reference's criterion for its formers is -55 dB, tests/test_api.py:151-167; with
    exact backprojection the gate alone is measured: -98 dB unwindowed, -110 dB windowed)."""
    col, S, grid, g = gated_scene
    win = oracle.window(col.Np, col.K) if window else None
    gated, full = _gate_images(col, S, grid, g, oracle.range_gate(S, g, win), win)
    e = db(gated, full)
    print(f'range gate, window {window}: {g["k1"] - g["k0"]} of {col.K} samples, {e:.1f} dB')
    assert e < -85


def test_range_gate_planted_mistakes(gated_scene):
    """Each plausible mistake in the gate fails the image check: the range cut kept as one block (no negative
    ranges), the transforms swapped, the Kp / N scale dropped."""
    col, S, grid, g = gated_scene
    npl, Kp, N, k0, k1 = g['npl'], g['Kp'], g['N'], g['k0'], g['k1']
    z = np.zeros((col.Np, Kp), np.complex128)
    z[:, npl:npl + col.K] = S
    q = np.fft.ifft(z, axis=1)
    h = N // 2
    folded = np.concatenate([q[:, :N - h], q[:, Kp - h:]], axis=1)
    wrong = {
        'one block': np.fft.fft(q[:, :N], axis=1)[:, k0:k1] * g['taper'],
        'swapped transforms': np.fft.ifft(np.fft.fft(z, axis=1)[:, list(range(N - h)) + list(range(Kp - h, Kp))], axis=1)[:, k0:k1] * g['taper'] * N * N / Kp,
        'no scale': np.fft.fft(folded, axis=1)[:, k0:k1] * g['taper'] * (N / Kp),
    }
    for name, Sg in wrong.items():
        gated, full = _gate_images(col, S, grid, g, Sg, None)
        e = db(gated, full)
        print(f'{name}: {e:.1f} dB')
        assert e > -20, name


def test_range_gate_plan():
    """The plan's shape: a narrow image of a wide-band history keeps few samples; the taper is 1 over the band and
    falls to 0 across the guards; an image that fills the swath is not worth gating (the reference's rule: gated
    only when N < Kp and fewer than half the samples are kept, fastsar/api.py:409-411)."""
    from fastsar_amd import gate
    col, _, grid = wide_band(ncl=1)
    g = gate.gate_plan(col.ant, None, col.fmin, col.df, col.K, **grid)
    kept = g['k1'] - g['k0']
    assert g['N'] < g['Kp'] and kept < col.K // 20 and g['N'] % 2 == 0
    f = g['fmin'] + g['df'] * np.arange(kept)
    band = (col.fmin, col.fmin + (col.K - 1) * col.df)
    inside = (f >= band[0]) & (f <= band[1])
    assert np.allclose(g['taper'][inside], g['Kp'] / g['N'], rtol=1e-6)
    assert np.all(g['taper'][~inside] < g['Kp'] / g['N']) and np.all(g['taper'] > 0)
    assert f[0] >= band[0] - 0.125 * (band[1] - band[0]) - 1e-3 and f[-1] <= band[1] + 0.125 * (band[1] - band[0]) + 1e-3
    small = sim.make_collect(res=0.5, scene=60.0, r0=5e3)
    gs = gate.gate_plan(small.ant, None, small.fmin, small.df, small.K, 256, 256, 0.25, 0.25, (1.0, 0, 0), (0, 1.0, 0))
    assert not (gs['N'] < gs['Kp'] and gs['k1'] - gs['k0'] < small.K // 2)
