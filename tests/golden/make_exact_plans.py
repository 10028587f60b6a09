"""Writes tests/golden/exact_plans.json: the planning decisions of ExactFormer on the test and benchmark geometries,
as fastsar_amd/exact.py at commit 6efa9ad makes them (that planner restates fastsar/exact.py:415-540 line by line).
The planner was rewritten afterwards; tests/test_exact_plan.py checks that the rewritten one makes the same
decisions: expansion tile, tile counts, nfft, window length W and every pulse's first bin exactly, the predicted
error, bin and phase scales and the tile centres to 1e-12 relative.

    python tests/golden/make_exact_plans.py [rev]      # rev: the git revision whose planner writes the table

Planning touches no GPU: the backend check is bypassed while the table is written.
"""
import hashlib
import importlib.util
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def cases():
    """(name, antenna track, fmin, df, K, grid, keywords) of every recorded plan"""
    from fastsar_amd import sim
    out = []
    g = dict(nx=100, ny=120, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    for r0 in (20e3, 600e3, 3e3, 500.0, 250.0):
        col = sim.make_collect(res=0.5, scene=60.0, r0=r0)
        out.append((f'make_collect r0 {r0:.0f} m 100x120', col, g, {}))
    col = sim.make_collect(res=0.5, scene=60.0, r0=20e3)
    out.append(('make_collect 20 km linear upsample 4 chunk 37', col, g, dict(interp='linear', upsample=4, chunk=37)))
    out.append(('make_collect 20 km centre offset, own ref', col, g,
                dict(center=(3.0, -2.0, 0.5), ref=np.linalg.norm(col.ant, axis=1) + 1.5)))
    c1 = sim.spotlight(512, 1024, 3.0)
    out.append(('C1 512^2', c1, dict(nx=512, ny=512, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0)), {}))
    c2 = sim.spotlight(16384, 8192, 6.0)
    sp = 6000.0 / 12288
    out.append(('C2 12288x8192', c2, dict(nx=12288, ny=8192, spx=sp, spy=sp, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0)), {}))
    out.append(('C2 geometry 4096^2 linear', c2, dict(nx=4096, ny=4096, spx=sp, spy=sp, e1=(0.0, 1.0, 0.0),
                                                      e2=(1.0, 0.0, 0.0)), dict(interp='linear')))
    sc = sim.uneven_batch(seed=5, n=48)
    for i in (0, 11, int(np.argmax([s['grid']['nx'] for s in sc]))):
        s = sc[i]
        out.append((f'C5 image {i}', s['col'], s['grid'], {}))
    return out


def load_exact(rev):
    src = subprocess.run(['git', '-C', ROOT, 'show', f'{rev}:fastsar_amd/exact.py'], check=True, capture_output=True,
                         text=True).stdout
    d = tempfile.mkdtemp()
    path = os.path.join(d, 'exact_at_rev.py')
    with open(path, 'w') as fh:
        fh.write(src)
    spec = importlib.util.spec_from_file_location('fastsar_amd._exact_at_rev', path)
    mod = importlib.util.module_from_spec(spec)
    mod.__package__ = 'fastsar_amd'
    spec.loader.exec_module(mod)
    return mod


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.int64).tobytes()).hexdigest()


def record(ex, col, grid, kw):
    f = ex.ExactFormer(col.ant, col.fmin, col.df, col.K, **grid, **kw)
    idx = np.linspace(0, len(f.blo) - 1, 9).astype(int)
    cidx = np.linspace(0, len(f.cen) - 1, 7).astype(int)
    return dict(tile=[int(f.tx), int(f.ty)], ntx=int(f.ntx), nty=int(f.nty), nfft=int(f.nfft), W=int(f.W),
                predicted_db=float(f.predicted_error_db), inv_dr=float(f.inv_dr), kcyc=float(f.kcyc),
                blo_sha256=digest(f.blo), blo_at=[[int(i), int(f.blo[i])] for i in idx],
                cen_at=[[int(i)] + [float(v) for v in f.cen[i]] for i in cidx])


def main():
    rev = sys.argv[1] if len(sys.argv) > 1 else '6efa9ad'
    import fastsar_amd.api as api
    api._backend = lambda backend: 'hip'           # planning only
    ex = load_exact(rev)
    table = dict(rev=rev, plans=[])
    for name, col, grid, kw in cases():
        print(name, flush=True)
        kws = {k: (v.tolist() if isinstance(v, np.ndarray) else v) for k, v in kw.items()}
        table['plans'].append(dict(name=name, grid=grid, kw=kws, **record(ex, col, grid, kw)))
    with open(os.path.join(HERE, 'exact_plans.json'), 'w') as fh:
        json.dump(table, fh, indent=0)


if __name__ == '__main__':
    main()
