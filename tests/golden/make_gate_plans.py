"""Writes tests/golden/gate_plans.json: the range-gate plans of fastsar_amd/gate.py at commit 51c78b4 (whose
gate_plan restates fastsar/api.py:272-300 line by line) on wide-band and benchmark geometries. gate_plan was
rewritten afterwards; tests/test_gate_plan.py checks that the rewritten one makes the same plans: npl, Kp, N, k0, k1
exactly, fmin and df to 1e-12 relative, the taper to float32 rounding.

    python tests/golden/make_gate_plans.py [rev]
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

C = 299792458.0


def wide_band(K, P, r0=30e3, res=0.5, graze=40.0):
    """A linear track at r0 whose K samples span a band far wider than a small image needs."""
    from fastsar_amd import sim
    df = C / (2 * res) / K
    y = (np.arange(P) - (P - 1) / 2) * (r0 * C / 9.6e9 / (2 * res) / P)
    psi = np.radians(graze)
    ant = np.stack([np.full(P, -r0 * np.cos(psi)), y, np.full(P, r0 * np.sin(psi))], 1)
    return sim.Collect(fmin=9.6e9 - (K // 2) * df, df=df, K=K, ant=ant, res=res)


def cases():
    """(name, collection, ref offset or None, grid) of every recorded plan"""
    from fastsar_amd import sim
    g = lambda nx, ny, sp, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0): dict(nx=nx, ny=ny, spx=sp, spy=sp, e1=e1, e2=e2)
    out = [('wide band 4096 x 300, 96^2 at 0.4 m', wide_band(4096, 300), None, g(96, 96, 0.4)),
           ('wide band 8192 x 257, 48^2 at 0.4 m', wide_band(8192, 257), None, g(48, 48, 0.4)),
           ('wide band 2000 x 600, 128 x 64 at 0.3 m, ref + 2.5 m', wide_band(2000, 600, r0=12e3), 2.5, g(128, 64, 0.3)),
           ('wide band 6000 x 1000 at 60 deg, 200^2 at 0.5 m, axes swapped', wide_band(6000, 1000, graze=60.0), None,
            g(200, 200, 0.5, (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)))]
    mc = sim.make_collect(res=0.5, scene=60.0, r0=5e3)
    out.append(('make_collect 5 km, 256^2 at 0.25 m', mc, None, g(256, 256, 0.25, (1.0, 0.0, 0.0), (0.0, 1.0, 0.0))))
    c2 = sim.spotlight(16384, 8192, 6.0)
    out.append(('C2 geometry, 512^2 at 0.49 m', c2, None, g(512, 512, 6000.0 / 12288)))
    return out


def load_gate(rev):
    src = subprocess.run(['git', '-C', ROOT, 'show', f'{rev}:fastsar_amd/gate.py'], check=True, capture_output=True,
                         text=True).stdout
    d = tempfile.mkdtemp()
    path = os.path.join(d, 'gate_at_rev.py')
    with open(path, 'w') as fh:
        fh.write(src)
    spec = importlib.util.spec_from_file_location('fastsar_amd._gate_at_rev', path)
    mod = importlib.util.module_from_spec(spec)
    mod.__package__ = 'fastsar_amd'
    spec.loader.exec_module(mod)
    return mod


def main():
    rev = sys.argv[1] if len(sys.argv) > 1 else '51c78b4'
    gm = load_gate(rev)
    table = dict(rev=rev, plans=[])
    for name, col, dref, grid in cases():
        ref = None if dref is None else np.linalg.norm(col.ant, axis=1) + dref
        p = gm.gate_plan(col.ant, ref, col.fmin, col.df, col.K, **grid)
        table['plans'].append(dict(name=name, npl=p['npl'], Kp=p['Kp'], N=p['N'], k0=p['k0'], k1=p['k1'],
                                   fmin=p['fmin'], df=p['df'], taper=[float(v) for v in p['taper']]))
        print(name, p['Kp'], p['N'], p['k1'] - p['k0'], flush=True)
    with open(os.path.join(HERE, 'gate_plans.json'), 'w') as fh:
        json.dump(table, fh, indent=0)


if __name__ == '__main__':
    main()
