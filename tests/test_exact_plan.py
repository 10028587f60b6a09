"""The exact former's host plan (ExactPlan) on the CPU: the same decisions as the planner that restated
fastsar/exact.py:415-540 (tests/golden/exact_plans.json, written by tests/golden/make_exact_plans.py), and the
properties that fix them: every pixel's interpolation taps fall inside its pulse's kept window (the reference's
tests/test_exact.py:163-169), the expansion tile is the first candidate within the phase limit, and a grid too near the
antenna or too deep in range is refused."""
import hashlib
import json
import os

import numpy as np
import pytest

from fastsar_amd import sim
from fastsar_amd.exact import PHASE_LIMIT, TILES, ExactPlan, expansion_tile, plane_points
from fastsar_amd.plan import C

HERE = os.path.dirname(os.path.abspath(__file__))
OLD_TILES = ((32, 32), (16, 16), (8, 16), (8, 8))       # the candidates of the recorded planner


def _table():
    with open(os.path.join(HERE, 'golden', 'exact_plans.json')) as fh:
        return json.load(fh)['plans']


def _collections():
    from tests.golden.make_exact_plans import cases
    return {name: (col, grid, kw) for name, col, grid, kw in cases()}


def _plan(col, grid, kw, tiles=TILES):
    ant = np.asarray(col.ant, np.float64)
    ref = np.asarray(kw['ref'], np.float64) if 'ref' in kw else np.linalg.norm(ant, axis=1)
    centre = np.asarray(kw.get('center', (0.0, 0.0, 0.0)), np.float64)
    return ExactPlan(ant, ref, col.fmin, col.df, col.K, grid['nx'], grid['ny'], grid['spx'], grid['spy'],
                     np.asarray(grid['e1'], np.float64), np.asarray(grid['e2'], np.float64), centre,
                     kw.get('upsample', 8), tiles)


@pytest.fixture(scope='module')
def collections():
    return _collections()


@pytest.mark.parametrize('i', range(len(_table())))
def test_plan_matches_recorded_decisions(collections, i):
    rec = _table()[i]
    col, grid, kw = collections[rec['name']]
    p = _plan(col, grid, kw, OLD_TILES)
    assert [p.tx, p.ty] == rec['tile'] and (p.ntx, p.nty) == (rec['ntx'], rec['nty'])
    assert p.nfft == rec['nfft'] and p.W == rec['W']
    assert hashlib.sha256(np.ascontiguousarray(p.blo, np.int64).tobytes()).hexdigest() == rec['blo_sha256']
    for j, v in rec['blo_at']:
        assert p.blo[j] == v
    assert p.predicted_error_db == pytest.approx(rec['predicted_db'], rel=1e-12)
    assert p.inv_dr == pytest.approx(rec['inv_dr'], rel=1e-12) and p.kcyc == pytest.approx(rec['kcyc'], rel=1e-12)
    for j, *xyz in rec['cen_at']:
        np.testing.assert_allclose(p.cen[j], xyz, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize('name', ['make_collect r0 20000 m 100x120', 'make_collect r0 500 m 100x120',
                                  'make_collect 20 km centre offset, own ref', 'make_collect 20 km linear upsample 4 chunk 37'])
def test_window_covers_every_pixel(collections, name):
    """brute force over every pixel and pulse: the cubic taps (bins floor(t) - 1 .. floor(t) + 2) of the exact range
    lie in [blo, blo + W), and the grid spans fewer bins than the profile's period"""
    col, grid, kw = collections[name]
    p = _plan(col, grid, kw)
    ref = np.asarray(kw['ref']) if 'ref' in kw else np.linalg.norm(col.ant, axis=1)
    pts = plane_points(grid['nx'], grid['ny'], grid['spx'], grid['spy'], grid['e1'], grid['e2']).reshape(-1, 3)
    pts = pts + np.asarray(kw.get('center', (0.0, 0.0, 0.0)))
    for p0 in range(0, len(col.ant), 64):
        a = col.ant[p0:p0 + 64]
        r = np.linalg.norm(pts[None] - a[:, None], axis=2)
        b = np.floor((r - ref[p0:p0 + 64, None]) * p.inv_dr).astype(np.int64) + p.nfft // 2
        assert (b - 1 >= p.blo[p0:p0 + 64, None]).all()
        assert (b + 2 < p.blo[p0:p0 + 64, None] + p.W).all()
    assert p.need <= p.nfft


def test_expansion_tile_rule():
    """the first candidate whose remainder phase meets the limit; past every candidate the smallest usable one; none
    when even the smallest tile is closer than four half-diagonals"""
    kmax = 4 * np.pi * 9.75e9 / C
    sp = 0.5
    for rmin in (20e3, 3e3, 500.0, 250.0, 120.0, 60.0):
        pick = expansion_tile(rmin, kmax, sp, sp)
        phases = []
        for tx, ty in TILES:
            h = 0.5 * np.hypot((tx - 1) * sp, (ty - 1) * sp)
            r = rmin - h
            phases.append(None if r < 4 * h else kmax * h ** 4 / (2 * r ** 3))
        ok = [i for i, ph in enumerate(phases) if ph is not None and ph <= PHASE_LIMIT]
        usable = [i for i, ph in enumerate(phases) if ph is not None]
        if ok:
            assert pick[:2] == TILES[ok[0]] and pick[3] == pytest.approx(phases[ok[0]], rel=1e-14)
        elif usable:
            assert pick[:2] == TILES[usable[-1]] and pick[3] > PHASE_LIMIT
        else:
            assert pick is None


def test_small_tiles_near_the_antenna(collections):
    """at 250 m the recorded planner's smallest tile (8 x 8) misses the limit; the 4 x 4 tile meets it"""
    col, grid, kw = collections['make_collect r0 250 m 100x120']
    old = _plan(col, grid, kw, OLD_TILES)
    new = _plan(col, grid, kw)
    assert (old.tx, old.ty) == (8, 8) and old.phase > PHASE_LIMIT
    assert (new.tx, new.ty) == (4, 4) and new.phase <= PHASE_LIMIT


def test_refusals():
    col = sim.make_collect(res=0.5, scene=60.0, r0=20e3)
    ant = col.ant
    ref = np.linalg.norm(ant, axis=1)
    e1, e2 = np.array([0.0, 1.0, 0.0]), np.array([1.0, 0.0, 0.0])
    # a grid far deeper in range than c / (2 df)
    p = ExactPlan(ant, ref, col.fmin, col.df, col.K, 40, 4000, 0.5, 0.5, e1, e2, np.zeros(3), 8)
    assert p.need > p.nfft
    # an antenna inside the grid: no usable tile
    near = ant * (1.0 / np.linalg.norm(ant, axis=1))[:, None] * 3.0
    with pytest.raises(ValueError, match='too close'):
        ExactPlan(near, np.linalg.norm(near, axis=1), col.fmin, col.df, col.K, 100, 100, 0.5, 0.5, e1, e2, np.zeros(3), 8)
