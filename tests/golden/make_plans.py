"""Writes tests/golden/plans.json: the decisions of the FFBP planner on the benchmark geometries, as the restatement
of fastsar/ffbp2.py:31-342 at commit 305a704 (fastsar_amd/plan.py there follows the reference line by line) makes
them. The planner was rewritten afterwards; tests/test_plan.py checks that the rewritten one makes the same
decisions: every integer (splits, decimations, margins, shapes, filter lengths and paddings) exactly, the pass bands,
frequencies and predicted errors to 1e-12 relative, the taps to 1e-12 of their largest, and the host levels'
rotations at sampled entries to float32 rounding.

    python tests/golden/make_plans.py [rev]      # rev: the git revision whose planner writes the table (305a704)
"""
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
    """(name, collection, grid, plan keywords) of every recorded plan"""
    from fastsar_amd import sim
    c1 = sim.spotlight(512, 1024, 3.0)
    g1 = dict(nx=512, ny=512, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    c2 = sim.spotlight(16384, 8192, 6.0)
    sp = 6000.0 / 12288
    g2 = dict(nx=12288, ny=8192, spx=sp, spy=sp, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    c4 = sim.spotlight(1024, 1024, 3.0)
    mc = sim.make_collect(res=0.5, scene=60.0, r0=5e3)
    g128 = dict(nx=128, ny=128, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    out = []
    for T in (32, 16, 8):
        out.append((f'C1 T{T} 3 levels', c1, g1, dict(T=T, nlev=3)))
    out.append(('C1 T8 4 levels 80 dB', c1, g1, dict(T=8, nlev=4, atten=80.0)))
    out.append(('C2 T32 3 levels', c2, g2, dict(T=32, nlev=3)))
    out.append(('C2 T16 4 levels', c2, g2, dict(T=16, nlev=4)))
    out.append(('C2 T8 4 levels', c2, g2, dict(T=8, nlev=4)))
    out.append(('C2 T8 4 levels 80 dB', c2, g2, dict(T=8, nlev=4, atten=80.0)))
    out.append(('C4 T16 3 levels', c4, g1, dict(T=16, nlev=3)))
    out.append(('make_collect 128 T16 2 levels', mc, g128, dict(T=16, nlev=2)))
    out.append(('make_collect 128 T8 3 levels pmax 0.3', mc, g128, dict(T=8, nlev=3, pmax=0.3)))
    sc = sim.uneven_batch(seed=5, n=48)
    for i in (0, 1, int(np.argmax([s['grid']['nx'] for s in sc]))):
        s = sc[i]
        out.append((f'C5 image {i} T8 4 levels 80 dB', s['col'], s['grid'], dict(T=8, nlev=4, atten=80.0)))
    return out


def load_planner(rev):
    src = subprocess.run(['git', '-C', ROOT, 'show', f'{rev}:fastsar_amd/plan.py'], check=True, capture_output=True,
                         text=True).stdout
    d = tempfile.mkdtemp()
    path = os.path.join(d, 'fastsar_amd_plan_at_rev.py')
    with open(path, 'w') as fh:
        fh.write(src)
    spec = importlib.util.spec_from_file_location('fastsar_amd_plan_at_rev', path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def record(pl, col, grid, kw):
    p = pl.make_plan(col, grid['nx'], grid['ny'], grid['spx'], grid['spy'], e1=grid['e1'], e2=grid['e2'], **kw)
    lv = []
    for l in p['levels']:
        e = {k: int(l[k]) for k in ('sx', 'sy', 'C', 'Dk', 'Dp', 'K', 'P', 'Ko', 'Po', 'mk', 'mp')}
        e.update(pass_k=float(l['pass_k']), pass_p=float(l['pass_p']), f0=float(l['f0']), df=float(l['df']),
                 pidx0=float(l['pidx'][0]), npidx=int(len(l['pidx'])))
        for ax in ('k', 'p'):
            fr = l[f'fir_{ax}']
            if fr is None:
                e[f'fir_{ax}'] = None
            else:
                e[f'fir_{ax}'] = dict(L=int(fr['L']), pl=int(fr['pl']), pr=int(fr['pr']), n_out=int(fr['n_out']),
                                      kern=[float(v) for v in fr['kern']])
        lv.append(e)
    f = p['final']
    fmax = col.fmin + col.K * col.df
    T = kw['T']
    err = pl.final_phase_error(col.ant, fmax, grid['nx'], grid['ny'], grid['spx'], grid['spy'], grid['e1'], grid['e2'], T)
    choice = pl.choose_T(col.ant, fmax, grid['nx'], grid['ny'], grid['spx'], grid['spy'], grid['e1'], grid['e2'], -40.0)
    coll = pl.collection_arrays(p, col.ant)
    rng = np.random.default_rng(0)
    rot = []
    for i in range(min(2, len(p['levels']))):
        c0, sl = coll['levels'][i]['c0'], coll['levels'][i]['slope']
        idx = [tuple(int(rng.integers(0, n)) for n in c0.shape) for _ in range(16)]
        rot.append(dict(idx=[list(t) for t in idx], c0=[float(c0[t]) for t in idx], slope=[float(sl[t]) for t in idx]))
    return dict(split=[list(map(int, s)) for s in p['split']], Nx=int(p['Nx']), Ny=int(p['Ny']), ox=int(p['ox']),
                oy=int(p['oy']), levels=lv, final=dict(P=int(f['P']), K=int(f['K']), f0=float(f['f0']), df=float(f['df']),
                                                      fc=float(f['fc']), ntiles=int(f['cen'].shape[0])),
                predicted_db=float(err), auto_T=int(choice[0]), auto_db=float(choice[1]), rotations=rot,
                final_r0=[float(v) for v in coll['final']['r0'][::max(1, len(coll['final']['r0']) // 8)]])


def main():
    rev = sys.argv[1] if len(sys.argv) > 1 else '305a704'
    pl = load_planner(rev)
    table = dict(rev=rev, plans=[])
    for name, col, grid, kw in cases():
        print(name, flush=True)
        table['plans'].append(dict(name=name, grid=grid, kw=kw, **record(pl, col, grid, kw)))
    with open(os.path.join(HERE, 'plans.json'), 'w') as fh:
        json.dump(table, fh, indent=0)


if __name__ == '__main__':
    main()
