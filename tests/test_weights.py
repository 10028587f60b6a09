"""Aperture weights on the host (fastsar_amd.api.final_weights): each final subaperture's weight at each tile
against the plain mean of the weight over the pulses that subaperture spans, and the gradient along e1 against a
weight whose slope is known. The thresholds are those of fastsar's own unit checks (tests/test_units.py:204-229)."""
import numpy as np
import pytest

from fastsar_amd import plan as pl, sim
from fastsar_amd.api import final_weights

GRID = dict(nx=128, ny=128, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
E1, E2 = np.array(GRID['e1']), np.array(GRID['e2'])


def _hann(P):
    return lambda p: 0.5 - 0.5 * np.cos(2 * np.pi * (np.asarray(p) + 0.5) / P)


def _wfun(P):
    hann = _hann(P)

    def wfun(q, p):
        q = np.asarray(q)
        return hann(p)[:, None] * (1.0 + 0.004 * (q @ E1) - 0.003 * (q @ E2))[None, :]
    return wfun


@pytest.fixture(scope='module')
def col():
    return sim.make_collect(res=0.5, scene=60.0, r0=5e3)


def _spans(plan):
    """(centre on the input pulse axis, width) of every final subaperture, composed through the levels by hand"""
    idx, D = np.arange(plan['final']['P'], dtype=np.float64), 1.0
    for lv in reversed(plan['levels']):
        m, Dp = lv['mp'], lv['Dp']
        idx = Dp * (idx - m) + (Dp - 1) / 2.0
        D *= Dp
    return idx, D


@pytest.mark.parametrize('nlev', [2, 3])
def test_final_weights_mean_over_pulses(col, nlev):
    P = col.Np
    wfun = _wfun(P)
    plan = pl.make_plan(col, 128, 128, 0.5, 0.5, T=16, nlev=nlev, e1=GRID['e1'], e2=GRID['e2'])
    wf = final_weights(plan, wfun, P)
    idx, D = _spans(plan)
    cen = plan['final']['cen']
    want = np.zeros_like(wf, dtype=np.float64)
    for f in range(len(idx)):
        p = np.arange(int(np.ceil(idx[f] - D / 2)), int(np.floor(idx[f] + D / 2)) + 1)
        p = p[(p >= 0) & (p < P)]
        if not len(p):
            p = np.array([int(np.clip(np.rint(idx[f]), 0, P - 1))])
        want[:, f] = wfun(cen, p).mean(0)
    err = float(np.abs(wf - want).max())
    print(f'{nlev} levels, subapertures of {D:.0f} pulses: max abs {err:.4f}')
    assert err <= 0.006
    g = final_weights(plan, wfun, P, grad=True)
    gx = _hann(P)(np.clip(np.rint(idx), 0, P - 1))[None, :] * 0.004
    assert float(np.abs(g[1] - gx).max() / 0.004) <= 0.006
    gy = -_hann(P)(np.clip(np.rint(idx), 0, P - 1))[None, :] * 0.003
    assert float(np.abs(g[2] - gy).max() / 0.003) <= 0.006
    assert np.array_equal(g[0], wf)


def test_final_weights_padded_span(col):
    """a collection of n pulses inside a longer plan (padded=): subapertures outside the collection take the nearest
    recorded pulse's weight, and those inside the same weights as the unpadded plan's quadrature gives"""
    P = col.Np
    plan = pl.make_plan(col, 128, 128, 0.5, 0.5, T=16, nlev=2, e1=GRID['e1'], e2=GRID['e2'])
    first, n = P // 4, P // 2

    def step(q, p):
        return np.repeat((np.asarray(p, np.float64) + 1.0)[:, None], len(q), 1)
    wf = final_weights(plan, step, n, first=first)
    idx, D = _spans(plan)
    lo, hi = first, first + n - 1
    outside = (idx + D / 2 < lo - 0.5) | (idx - D / 2 > hi + 0.5)
    near = np.clip(np.rint(idx), lo, hi) + 1.0
    assert np.allclose(wf[:, outside], near[None, outside], rtol=1e-6)
    inside = (idx - D / 2 >= lo - 0.5) & (idx + D / 2 <= hi + 0.5)
    # the quadrature nodes are whole pulses: a linear weight comes out within half a pulse's step of its mean
    assert np.abs(wf[:, inside] - (idx[inside] + 1.0)[None, :]).max() <= 0.5
