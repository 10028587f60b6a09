"""The range gate's host plan against the plans recorded from the restatement of fastsar/api.py:272-300
(tests/golden/gate_plans.json, written by tests/golden/make_gate_plans.py): every integer exactly, the gated
history's first frequency and spacing to 1e-12 relative, the taper to float32 rounding. The plan's properties are
pinned in tests/test_oracle.py (test_range_gate_plan) and its images against float64 backprojection there and in
tests/test_gate_gpu.py."""
import json
import os

import numpy as np
import pytest

from fastsar_amd import gate

HERE = os.path.dirname(os.path.abspath(__file__))


def _table():
    with open(os.path.join(HERE, 'golden', 'gate_plans.json')) as fh:
        return json.load(fh)['plans']


@pytest.fixture(scope='module')
def cases():
    from tests.golden.make_gate_plans import cases as c
    return {name: rest for name, *rest in c()}


@pytest.mark.parametrize('i', range(len(_table())))
def test_gate_plan_matches_recorded(cases, i):
    rec = _table()[i]
    col, dref, grid = cases[rec['name']]
    ref = None if dref is None else np.linalg.norm(col.ant, axis=1) + dref
    g = gate.gate_plan(col.ant, ref, col.fmin, col.df, col.K, **grid)
    assert (g['npl'], g['Kp'], g['N'], g['k0'], g['k1']) == (rec['npl'], rec['Kp'], rec['N'], rec['k0'], rec['k1'])
    assert g['fmin'] == pytest.approx(rec['fmin'], rel=1e-12) and g['df'] == pytest.approx(rec['df'], rel=1e-12)
    np.testing.assert_allclose(g['taper'], np.asarray(rec['taper'], np.float32), rtol=2e-7, atol=1e-7)


def test_taper_shape():
    """1 on the band, cos^2 of the fraction of the guard crossed times pi / 2 within it, 0 beyond"""
    band, guard = (10.0, 20.0), 2.0
    f = np.array([9.0, 10.0, 15.0, 20.0, 20.5, 21.0, 21.5, 22.0, 23.0, 7.5])
    want = [np.cos(np.pi / 4) ** 2, 1, 1, 1, np.cos(np.pi / 8) ** 2, np.cos(np.pi / 4) ** 2, np.cos(3 * np.pi / 8) ** 2,
            0, 0, 0]
    np.testing.assert_allclose(gate.taper(f, band, guard), want, atol=1e-15)
