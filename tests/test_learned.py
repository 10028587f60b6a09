"""The learned decimators' host side and the oracle of their training, on the CPU: the network's gradients and the
least-squares loss of the oracle pinned by central finite differences and by the direct sum over pairs, and the
fallback filter's shape."""
import numpy as np

import oracle
from fastsar_amd import learned


def _net(T, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(learned.nparam(T)) * 0.3


def _stages(rng, T, Ds=(3, 2)):
    st, off = [], 0
    for D in Ds:
        L = D * T
        X = rng.standard_normal((400, L)) + 1j * rng.standard_normal((400, L))
        y = X @ rng.standard_normal(L) * 0.3 + 0.05 * (rng.standard_normal(400) + 1j * rng.standard_normal(400))
        R = (X.conj().T @ X).real
        v = (X.conj().T @ y).real
        E = float(np.sum(np.abs(y) ** 2))
        st.append((D, L, off, R, v, E, rng.standard_normal(L) * 0.1, X, y))
        off += D
    return st, off


def test_oracle_gradients_by_finite_differences():
    T = 4
    rng = np.random.default_rng(1)
    st, npts = _stages(rng, T)
    pts = np.stack([(np.arange(npts) + 0.5) / 3, rng.uniform(0, 1, npts)], 1)
    p = _net(T)
    stages = [s[:7] for s in st]

    def total(q):
        out, _, _ = oracle.tapnet_forward(q, T, pts)
        return oracle.ls_loss_grad(out, T, stages)[0].sum()
    out, _, _ = oracle.tapnet_forward(p, T, pts)
    loss, dout = oracle.ls_loss_grad(out, T, stages)
    g = oracle.tapnet_backward(p, T, pts, dout)
    idx = rng.choice(len(p), 60, replace=False)
    eps = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = eps
        fd = (total(p + e) - total(p - e)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])
    # the loss from the statistics equals the direct sum over the pairs
    for (D, L, off, R, v, E, base, X, y), l in zip(st, loss):
        r = np.arange(L)
        h = base + out[off + r % D, r // D]
        direct = np.sum(np.abs(X @ h - y) ** 2) / E
        assert abs(direct - l) <= 1e-9 * max(1.0, direct)


def test_fallback_filter_shape():
    """the same-length Kaiser fallback: each phase sums to one, centred on the output, symmetric when D (T - 1) is
    even"""
    for D, T in ((7, 6), (2, 4), (3, 5), (11, 4)):
        h = learned.kaiser_same_length(D, T, 0.6, 3)
        L, pl = learned.placement(D, 3, T)
        assert len(h) == L == D * T
        for ph in range(D):
            assert abs(h[ph::D].sum() - 1) < 1e-12
        if (D * (T - 1)) % 2 == 0:
            assert np.allclose(h, h[::-1], atol=1e-12)
