"""The host planner (fastsar_amd/plan.py) against what fixes it independently of its own code: the plan shapes and
predicted errors derived by hand from the reference's planner (fastsar/ffbp2.py:192-307) for the benchmark
configurations, the closed-form properties of the decimation filters (fastsar/ffbp2.py:109-137), and brute-force
float64 geometry for the rotation phases (fastsar/ffbp2.py:313-342, 578-590)."""
import math

import numpy as np
import pytest

from fastsar_amd import plan as pl, sim

C1 = dict(P=512, K=1024, aperture=3.0, n=(512, 512), sp=0.5)
C2 = dict(P=16384, K=8192, aperture=6.0, n=(12288, 8192), sp=6000.0 / 12288)


def _plan(cfg, T, nlev, atten=70.0):
    col = sim.spotlight(cfg['P'], cfg['K'], cfg['aperture'])
    nx, ny = cfg['n']
    return col, pl.make_plan(col, nx, ny, cfg['sp'], cfg['sp'], T=T, nlev=nlev, atten=atten, e1=(0.0, 1.0, 0.0),
                             e2=(1.0, 0.0, 0.0))


# per level: (split x, split y, Dk, Dp, children so far, Po, Ko); then (final Pf, Qf, tiles). Derived by hand from
# fastsar/ffbp2.py:192-250 with pmax = 0.4 and 70 dB filters.
TABLE = [
    (C1, 32, 3, [(4, 4, 3, 1, 16, 512, 350), (2, 2, 2, 2, 64, 266, 183), (2, 2, 2, 2, 256, 143, 100)], (143, 100, 256)),
    (C1, 16, 3, [(8, 8, 7, 2, 64, 266, 155), (2, 2, 2, 2, 256, 143, 88), (2, 2, 2, 2, 1024, 82, 54)], (82, 54, 1024)),
    (C1, 8, 3, [(8, 8, 7, 2, 64, 266, 155), (4, 4, 4, 4, 1024, 75, 47), (2, 2, 2, 2, 4096, 48, 34)], (48, 34, 4096)),
    (C2, 32, 3, [(8, 8, 3, 1, 64, 16384, 2739), (8, 8, 9, 11, 4096, 1498, 313), (6, 4, 4, 6, 98304, 258, 87)],
     (258, 87, 98304)),
    (C2, 16, 4, [(8, 8, 3, 1, 64, 16384, 2739), (8, 8, 9, 11, 4096, 1498, 313), (6, 4, 4, 6, 98304, 258, 87),
                 (2, 2, 2, 2, 393216, 139, 54)], (139, 54, 393216)),
    (C2, 8, 4, [(8, 8, 3, 1, 64, 16384, 2739), (8, 8, 9, 11, 4096, 1498, 313), (8, 8, 8, 8, 262144, 196, 48),
                (3, 2, 2, 3, 1572864, 74, 34)], (74, 34, 1572864)),
]


@pytest.mark.parametrize('row', range(len(TABLE)), ids=[f'{"C1" if t[0] is C1 else "C2"}-T{t[1]}' for t in TABLE])
def test_plan_shapes(row):
    cfg, T, nlev, levels, final = TABLE[row]
    _, p = _plan(cfg, T, nlev)
    n = 1
    got = []
    for lv in p['levels']:
        n *= lv['C']
        got.append((lv['sx'], lv['sy'], lv['Dk'], lv['Dp'], n, lv['Po'], lv['Ko']))
    assert got == levels
    assert (p['final']['P'], p['final']['K'], p['final']['cen'].shape[0]) == final
    # the padded grid is the product of the splits times T on both axes, and covers the image
    sx = math.prod(lv['sx'] for lv in p['levels'])
    sy = math.prod(lv['sy'] for lv in p['levels'])
    assert sx * T >= cfg['n'][0] and sy * T >= cfg['n'][1]


def test_splits_examples():
    """choose_splits on the axes of the benchmark grids (fastsar/ffbp2.py:34-62): the smallest product of factors
    up to 8 that covers the axis, larger factors first"""
    assert pl.choose_splits(512, 32, 3) == (4, 2, 2)
    assert pl.choose_splits(12288, 32, 3) == (8, 8, 6)
    assert pl.choose_splits(8192, 32, 3) == (8, 8, 4)
    for n in range(1, 3000, 37):
        for T in (8, 16, 32):
            for nlev in (2, 3, 4):
                need = -(-n // T)
                if need > 16 ** nlev:
                    with pytest.raises(ValueError):
                        pl.choose_splits(n, T, nlev)
                    continue
                s = pl.choose_splits(n, T, nlev)
                assert len(s) == nlev and math.prod(s) >= need and list(s) == sorted(s, reverse=True)
                assert min(s) >= 2
                prods = {1}
                for _ in range(nlev):
                    prods = {a * f for a in prods for f in range(2, 9)}
                if max(s) <= 8:
                    # no product of nlev factors 2 .. 8 lies in [need, prod(s))
                    assert not any(need <= v < math.prod(s) for v in prods)
                else:
                    assert max(prods) < need


@pytest.mark.parametrize('cfg,expect', [(C1, (-39.5, -51.5, -63.6)), (C2, (-31.5, -43.6, -55.6))])
def test_predicted_final_error(cfg, expect):
    """final_phase_error at T = 32, 16, 8 against the values derived by hand from fastsar/ffbp2.py:276-297"""
    col = sim.spotlight(cfg['P'], cfg['K'], cfg['aperture'])
    nx, ny = cfg['n']
    fmax = col.fmin + (col.K - 1) * col.df
    got = [pl.final_phase_error(col.ant, fmax, nx, ny, cfg['sp'], cfg['sp'], (0.0, 1.0, 0.0), (1.0, 0.0, 0.0), T)
           for T in (32, 16, 8)]
    assert np.allclose(got, expect, atol=0.1), got
    # the residual phase grows with the square of the tile size: 12 dB per halving
    assert np.allclose(np.diff(got), -12.0, atol=0.5)
    # the default target (-40 dB) takes T = 16 on both grids: 32 misses it
    T, e = pl.choose_T(col.ant, fmax, nx, ny, cfg['sp'], cfg['sp'], (0.0, 1.0, 0.0), (1.0, 0.0, 0.0), target_db=-40.0)
    assert T == 16 and abs(e - got[1]) < 1e-9


@pytest.mark.parametrize('n_in,D,pf,atten', [(8192, 3, 0.95, 70.0), (2741, 9, 0.6, 70.0), (315, 8, 0.6, 80.0),
                                             (50, 2, 0.6, 80.0), (1000, 4, 0.8, 70.0), (1000, 7, 0.5, 60.0)])
def test_decimator_closed_form(n_in, D, pf, atten):
    """Kaiser-windowed sinc decimator (fastsar/ffbp2.py:109-137): every input's weights sum to one (each polyphase
    component of the kernel sums to 1), the passband is flat and the stopband below the design attenuation, the
    output count is ceil(n_in / D) + 2m, and output j is centred on input D (j - m) + (D - 1) / 2"""
    fr, m = pl.decimator_fir(n_in, D, pf, atten)
    k, L = fr['kern'], fr['L']
    assert len(k) == L and fr['n_out'] == -(-n_in // D) + 2 * m
    assert np.allclose([k[ph::D].sum() for ph in range(D)], 1.0, atol=1e-12)
    w = np.linspace(0, np.pi, 4001)
    H = np.abs(np.exp(-1j * np.outer(w, np.arange(L))) @ k) / D
    assert np.abs(H[w <= pf * np.pi / D] - 1).max() < 10 ** (-(atten - 10) / 20)
    assert 20 * np.log10(H[w >= (2 - pf) * np.pi / D].max()) < -(atten - 4)
    # symmetric kernel: its centre (L - 1) / 2 sits at input D (j - m) + (D - 1) / 2 for output j
    assert np.allclose(k, k[::-1], atol=1e-15)
    assert fr['pl'] == D * m - (D - 1) / 2 + (L - 1) / 2
    # the dense matrix (column j = the kernel placed at output j) has unit row sums away from the edges
    F, _ = pl.decimator_matrix(n_in, D, pf, atten) if n_in <= 2741 else (None, None)
    if F is not None:
        assert np.allclose(F.sum(1), 1.0, atol=1e-12)


def test_collection_arrays_bruteforce():
    """c0 and slope of the host levels against |cen - a| - |ref - a| evaluated directly in float64"""
    col, p = _plan(C1, 16, 3)
    coll = pl.collection_arrays(p, col.ant)
    ant = np.asarray(col.ant, np.float64)
    for i in range(pl.HOST_LEVELS):
        lv, e = p['levels'][i], coll['levels'][i]
        k0, k1 = pl.level_phase_constants(lv)
        for b in range(min(3, lv['ref'].shape[0])):
            for c in range(0, lv['C'], max(1, lv['C'] // 5)):
                ddr = np.array([math.dist(lv['cen'][b][c], a) - math.dist(lv['ref'][b], a) for a in ant])
                cyc = k0 * ddr
                dphi = np.abs((e['c0'][b, c] - (cyc - np.round(cyc)) + 0.5) % 1.0 - 0.5)
                assert dphi.max() < 1e-6
                assert np.allclose(e['slope'][b, c], k1 * ddr, rtol=1e-6, atol=1e-12)
        ant = pl.positions(ant, lv['pidx'])


def test_device_phases_bruteforce():
    """the cancellation-free rotation of the later levels against the plain difference of ranges in float64"""
    rng = np.random.default_rng(3)
    P, Cn, Np = 64, 5, 3
    a = np.stack([np.full(P, -8660.0), np.linspace(-200, 200, P), np.full(P, 5000.0)], 1)
    r0 = np.linalg.norm(a, axis=1)
    u = a / r0[:, None]
    refs = rng.uniform(-300, 300, (Np, 3))
    d = rng.uniform(-20, 20, (Cn, 3))
    k0c, k1 = 64.0, 0.0025
    c0, sl = pl.device_phases(u, r0, d, refs, k0c, k1)
    for n in range(Np):
        for c in range(Cn):
            ddr = np.linalg.norm(refs[n] + d[c] - a, axis=1) - np.linalg.norm(refs[n] - a, axis=1)
            cyc = ddr * k0c
            assert np.abs((c0[n, c] - (cyc - np.rint(cyc)) + 0.5) % 1.0 - 0.5).max() < 1e-6
            assert np.allclose(sl[n, c], ddr * k1, rtol=1e-6, atol=1e-12)


def test_tile_digits_cover_grid():
    """every final tile has its own place, and the tiles tile the padded grid exactly"""
    for T, nlev in ((32, 3), (8, 3)):
        _, p = _plan(C1, T, nlev)
        xy = pl.tile_digits(p)
        Nx = math.prod(lv['sx'] for lv in p['levels']) * T
        Ny = math.prod(lv['sy'] for lv in p['levels']) * T
        assert len({tuple(v) for v in xy}) == len(xy) == (Nx // T) * (Ny // T)
        assert xy[:, 0].min() == 0 and xy[:, 1].min() == 0 and xy[:, 0].max() == Nx - T and xy[:, 1].max() == Ny - T
        assert np.all(xy % T == 0)


def _golden():
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'plans.json')) as fh:
        return json.load(fh)


GOLDEN = _golden()


@pytest.mark.parametrize('case', GOLDEN['plans'], ids=[c['name'] for c in GOLDEN['plans']])
def test_plan_against_golden_table(case):
    """the planner's decisions against the committed table (tests/golden/make_plans.py wrote it with the line-by-line
    restatement of fastsar/ffbp2.py): every integer exactly, pass bands, frequencies and predicted errors to 1e-12,
    the taps to 1e-12 of their largest, sampled host rotations to float32 rounding"""
    from tests.golden.make_plans import cases
    col, grid, kw = {c[0]: c[1:] for c in cases()}[case['name']]
    p = pl.make_plan(col, grid['nx'], grid['ny'], grid['spx'], grid['spy'], e1=grid['e1'], e2=grid['e2'], **kw)
    assert [list(s) for s in p['split']] == case['split']
    assert (p['Nx'], p['Ny'], p['ox'], p['oy']) == (case['Nx'], case['Ny'], case['ox'], case['oy'])
    for lv, want in zip(p['levels'], case['levels'], strict=True):
        for k in ('sx', 'sy', 'C', 'Dk', 'Dp', 'K', 'P', 'Ko', 'Po', 'mk', 'mp'):
            assert lv[k] == want[k], (k, lv[k], want[k])
        assert len(lv['pidx']) == want['npidx'] and lv['pidx'][0] == want['pidx0']
        for k in ('pass_k', 'pass_p', 'f0', 'df'):
            assert lv[k] == pytest.approx(want[k], rel=1e-12, abs=1e-300), k
        for ax in ('k', 'p'):
            got, w = lv[f'fir_{ax}'], want[f'fir_{ax}']
            assert (got is None) == (w is None)
            if got is not None:
                assert (got['L'], got['pl'], got['pr'], got['n_out']) == (w['L'], w['pl'], w['pr'], w['n_out'])
                kw_ = np.asarray(w['kern'])
                assert np.abs(got['kern'] - kw_).max() <= 1e-12 * np.abs(kw_).max()
    f = p['final']
    assert (f['P'], f['K'], f['cen'].shape[0]) == (case['final']['P'], case['final']['K'], case['final']['ntiles'])
    for k in ('f0', 'df', 'fc'):
        assert f[k] == pytest.approx(case['final'][k], rel=1e-12)
    fmax = col.fmin + col.K * col.df
    args = (col.ant, fmax, grid['nx'], grid['ny'], grid['spx'], grid['spy'], grid['e1'], grid['e2'])
    assert pl.final_phase_error(*args, kw['T']) == pytest.approx(case['predicted_db'], abs=1e-9)
    T, e = pl.choose_T(*args, -40.0)
    assert T == case['auto_T'] and e == pytest.approx(case['auto_db'], abs=1e-9)
    coll = pl.collection_arrays(p, col.ant)
    for i, rot in enumerate(case['rotations']):
        c0, sl = coll['levels'][i]['c0'], coll['levels'][i]['slope']
        for t, a, b in zip(rot['idx'], rot['c0'], rot['slope']):
            t = tuple(t)
            assert abs((float(c0[t]) - a + 0.5) % 1.0 - 0.5) <= 2.0 ** -24 and float(sl[t]) == pytest.approx(b, rel=2.0 ** -23)
    r0 = coll['final']['r0']
    assert np.allclose(r0[::max(1, len(r0) // 8)], case['final_r0'], rtol=1e-15, atol=0)
