"""The learned decimators' HIP kernels against the float64 oracle (the least-squares statistics of a stage, the
network's forward and backward passes, the loss gradient), and what the trained filters gain: on the small spotlight
geometry at 1024 pulses, trained on scenes the method itself forms, images of held-out scenes against the FP64
oracle with the learned filters, the same-length Kaiser fallback and the plan's Kaiser design."""
import ctypes

import numpy as np
import pytest

import oracle
from fastsar_amd import _hip, learned, plan as pl, sim
from fastsar_amd.exact import simulate_device, bp_points_f64_device

pytestmark = pytest.mark.gpu


def _windows_numpy(x, y, D, pl_, L, axis):
    """pairs of a stage by direct indexing: x [B, P, K] (range: along K; pulse: along P) -> (X [n, L], Y [n])"""
    if axis == 'pulse':
        x, y = np.swapaxes(x, 1, 2), np.swapaxes(y, 1, 2)
    B, M, n_in = x.shape
    n_out = y.shape[2]
    rows = x.reshape(B * M, n_in)
    X = np.zeros((B * M, n_out, L), np.complex128)
    for j in range(n_out):
        for r in range(L):
            k = D * j + r - pl_
            if 0 <= k < n_in:
                X[:, j, r] = rows[:, k]
    return X.reshape(-1, L), y.reshape(-1).astype(np.complex128)


@pytest.mark.parametrize('axis', ['range', 'pulse'])
def test_ls_stats_vs_numpy(axis):
    rng = np.random.default_rng(3)
    B, P, K, D, T = 3, 37, 53, 3, 4
    L, pl_ = learned.placement(D, 2, T)
    if axis == 'range':
        n_in, n_out = K, -(-K // D) + 4
        x = (rng.standard_normal((B, P, K)) + 1j * rng.standard_normal((B, P, K))).astype(np.complex64)
        y = (rng.standard_normal((B, P, n_out)) + 1j * rng.standard_normal((B, P, n_out))).astype(np.complex64)
        args = (B * P, 1, K, 0, 1, n_out, 0, 1, K, n_out)
    else:
        n_in, n_out = P, -(-P // D) + 4
        x = (rng.standard_normal((B, P, K)) + 1j * rng.standard_normal((B, P, K))).astype(np.complex64)
        y = (rng.standard_normal((B, n_out, K)) + 1j * rng.standard_normal((B, n_out, K))).astype(np.complex64)
        args = (B * K, K, P * K, 1, K, n_out * K, 1, K, P, n_out)
    R, v, E = (_hip.DeviceArray(s, np.float64).zero() for s in ((L, L), (L,), (1,)))
    L_ = learned._bind()
    dx, dy = _hip.DeviceArray.from_host(x), _hip.DeviceArray.from_host(y)      # held until the kernel has run
    a = list(args)
    a.insert(9, 0)                                   # j0: every output
    _hip.check(L_.fsh_ls_stats(_hip.ptr(dx), _hip.ptr(dy), *a, D, pl_, L, _hip.ptr(R), _hip.ptr(v), _hip.ptr(E), None))
    X, Y = _windows_numpy(x, y, D, pl_, L, axis)
    Rn, vn, En = (X.conj().T @ X).real, (X.conj().T @ Y).real, np.sum(np.abs(Y) ** 2)
    assert np.allclose(R.to_host(), Rn, rtol=1e-10, atol=1e-9 * np.abs(Rn).max())
    assert np.allclose(v.to_host(), vn, rtol=1e-10, atol=1e-9 * np.abs(vn).max())
    assert abs(E.to_host()[0] - En) < 1e-10 * En


def test_tapnet_kernels_vs_oracle():
    """forward, loss gradient and backward on the device against the float64 oracle (float32 arithmetic)"""
    rng = np.random.default_rng(5)
    T = 6
    p = (rng.standard_normal(learned.nparam(T)) * 0.3).astype(np.float32)
    net = learned.TapNet(T, params=p)
    Ds = (7, 2, 3)
    pts = np.concatenate([np.stack([(np.arange(D) + 0.5) / D, np.full(D, s / 2)], 1) for s, D in enumerate(Ds)]).astype(np.float32)
    dpts = _hip.DeviceArray.from_host(pts)
    out, act = net.forward(dpts)
    want, _, _ = oracle.tapnet_forward(p, T, pts)
    got = out.to_host()
    assert np.allclose(got, want, rtol=1e-4, atol=1e-5 * np.abs(want).max())
    # a problem of three stages with random statistics
    stages, stats, off = [], {}, 0
    for i, D in enumerate(Ds):
        L = D * T
        X = rng.standard_normal((300, L)) + 1j * rng.standard_normal((300, L))
        y = X @ rng.standard_normal(L) * 0.2 + 0.1 * (rng.standard_normal(300) + 1j * rng.standard_normal(300))
        R, v, E = (X.conj().T @ X).real, (X.conj().T @ y).real, float(np.sum(np.abs(y) ** 2))
        stats[(i, 'range')] = dict(D=D, m=2, pass_frac=0.6, L=L, pl=0, pairs=300, R=R, v=v, E=E)
        stages.append((D, L, off, R, v, E, learned.kaiser_same_length(D, T, 0.6, 2)))
        off += D
    prob = learned.Problem(stats, 'range', 3, T)
    loss, dout = prob.loss_grad(out)
    lw, dw = oracle.ls_loss_grad(got, T, stages, wt=np.ones(3) / 3)
    assert np.allclose(loss.to_host(), lw, rtol=1e-9)
    assert np.allclose(dout.to_host(), dw, rtol=1e-5, atol=1e-6 * np.abs(dw).max())
    grad = net.backward(dpts, act, dout).to_host()
    gw = oracle.tapnet_backward(p, T, pts, dout.to_host())
    assert np.allclose(grad, gw, rtol=1e-3, atol=1e-5 * np.abs(gw).max())


def _scene(seed, P=1024):
    """the small spotlight geometry (300 MHz, 3 degrees, 10 km) at 1024 pulses x 1024 samples, 25 random targets
    and -30 dB clutter"""
    rng = np.random.default_rng(seed)
    col = sim.spotlight(P, 1024, 3.0)
    pos, amp = sim.point_scene(rng, 25, 256.0, 256.0)
    return col, pos, amp, simulate_device(col, pos, amp, clutter_sigma=sim.clutter_sigma(amp, -30.0), seed=seed)


@pytest.mark.parametrize('T', [6, 8])
def test_learned_filters_gain(T):
    """trained on 4 scenes, evaluated on 2 held-out ones: the learned filters (T taps per phase) against the Kaiser
    fallback of the same length and the plan's 70 dB Kaiser design"""
    from fastsar_amd.api import _window
    from fastsar_amd.ffbp_hip import make_ffbp_hip
    grid = dict(nx=512, ny=512, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    st = learned.Stats(T)
    plan = None
    for seed in range(4):
        col, pos, amp, dS = _scene(100 + seed)
        if plan is None:
            plan = pl.make_plan(col, grid['nx'], grid['ny'], grid['spx'], grid['spy'], T=16, nlev=3, e1=grid['e1'], e2=grid['e2'])
            coll = pl.collection_arrays(plan, col.ant)
            wp, wk = _window(col.Np, col.K)
        st.add_scene(plan, coll, dS, window=(wp, wk))
    print('stages', {k: (v['D'], v['L'], v['pairs']) for k, v in st.stages.items()}, flush=True)
    taps = learned.train(st, iters=2000, lr=2e-3, log=print)
    print('training (mean relative error, dB):', taps.history, flush=True)
    # held-out scenes: image error against the FP64 checker
    same = learned.LearnedTaps(T, 3, {})                       # no networks: the same-length fallback
    forms = dict(kaiser70=plan, fallback=same.apply(plan), learned=taps.apply(plan))
    res = {k: [] for k in forms}
    pts = oracle.plane_points(**grid).reshape(-1, 3)
    for seed in (900, 901):
        col, pos, amp, dS = _scene(seed)
        nfft = 1 << int(np.ceil(np.log2(64 * col.K)))
        ref = bp_points_f64_device(dS, col.ant, np.linalg.norm(col.ant, axis=1), col.fmin, col.df, pts, nfft, True, wp, wk,
                                   chunk=64).reshape(grid['nx'], grid['ny'])
        for k, p in forms.items():
            img = make_ffbp_hip(p, coll, window=(wp, wk))(dS)
            res[k].append(oracle.rel_db(img, ref))
    taps_per = {k: [lv['fir_k']['L'] if lv['fir_k'] else 1 for lv in p['levels']] for k, p in forms.items()}
    for k in forms:
        print(f'  {k:9s} range taps {taps_per[k]}: {np.mean(res[k]):.1f} dB')
    # the gain over the same-length fallback: 11 to 15 dB at 6 taps per phase, 2 to 3.4 dB at 8 (whose fallback is
    # already near the 70 dB design)
    assert np.mean(res['learned']) < np.mean(res['fallback']) - {6: 6.0, 8: 1.0}[T]
