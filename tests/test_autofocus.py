"""Phase gradient autofocus in the image domain, on the host (fastsar_amd.autofocus.pga; the published method the
reference's fastsar/autofocus.py:56-102 implements). Pinned by what the method must do: a smooth phase error put
on the azimuth spectrum of a focused, band-limited image is recovered up to its constant and linear terms (which
PGA cannot see), and the image comes back; a focused image is left nearly alone."""
import numpy as np
import pytest

from fastsar_amd import autofocus as af


def rel_db(a, b):
    g = np.vdot(a, b) / np.vdot(a, a)
    return 10 * np.log10(np.sum(np.abs(g * a - b) ** 2) / np.sum(np.abs(b) ** 2))


def scene(nx=128, ny=96, band=0.8, seed=7):
    """A focused image with azimuth along axis 0: bright points over weak clutter, band-limited along azimuth to
    the centred bins |m| < band nx / 2 -> (image, centred bin numbers m, in-band mask)."""
    rng = np.random.default_rng(seed)
    g = 0.05 * (rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny)))
    idx = rng.choice(nx * ny, 60, replace=False)
    g.flat[idx] += rng.uniform(2, 10, 60) * np.exp(2j * np.pi * rng.random(60))
    m = np.arange(nx) - nx // 2
    inband = np.abs(m) < band * nx / 2
    G = np.fft.fft(g, axis=0) * np.fft.ifftshift(inband)[:, None]
    return np.fft.ifft(G, axis=0), m, inband


def with_error(g, e):
    """g with exp(j e) on its azimuth spectrum, e [nx] in centred bin order."""
    return np.fft.ifft(np.fft.fft(g, axis=0) * np.exp(1j * np.fft.ifftshift(e))[:, None], axis=0)


@pytest.mark.parametrize('method', ['ml', 'lumv'])
@pytest.mark.parametrize('kind', ['quadratic', 'cubic', 'smooth random'])
def test_pga_recovers_injected_error(method, kind):
    g0, m, inband = scene()
    v = m / (0.8 * len(m) / 2)
    rng = np.random.default_rng(11)
    if kind == 'quadratic':
        e = 6.0 * v ** 2
    elif kind == 'cubic':
        e = 4.0 * v ** 3 - 3.0 * v ** 2
    else:
        w = np.convolve(rng.standard_normal(len(m) + 24), np.hanning(25) / np.hanning(25).sum(), 'valid')[:len(m)]
        e = 3.0 * w / np.abs(w[inband]).max()
    e = np.where(inband, e, 0.0)
    g = with_error(g0, e)
    img, phi, hist = af.pga(g, method=method)
    r = (e - phi)[inband]
    mi = m[inband].astype(np.float64)
    A = np.stack([np.ones_like(mi), mi], 1)
    line = A @ np.linalg.lstsq(A, r, rcond=None)[0]
    assert np.std(r - line) < 0.05, (kind, method, np.std(r - line))
    # what PGA left is a constant and a linear phase: the focused image shifted and turned
    lin = np.zeros(len(m))
    lin[inband] = line
    assert rel_db(img, with_error(g0, lin)) < -30
    assert rel_db(g, g0) > -10 > rel_db(img, with_error(g0, lin))
    assert hist[-1]['rms'] < 0.005 or len(hist) == 20


def test_pga_leaves_focused_image():
    g0, _, _ = scene(seed=9)
    img, phi, hist = af.pga(g0)
    assert rel_db(img, g0) < -30 and len(hist) <= 20


def test_pga_errors():
    with pytest.raises(ValueError):
        af.pga(np.zeros((16, 16), np.complex64), method='nope')
    with pytest.raises(ValueError):
        af.pga(np.zeros(16, np.complex64))


def test_pulse_bins_broadside():
    """A broadside track along e1: pulse p at along-track y_p sees the scene centre turned by y_p / |a_p|, a plane
    wave of 2 fc y_p / (c |a_p|) cycles/m along e1, bin -that x spx nx; the aperture centre is at DC."""
    from fastsar_amd.sim import C
    y = np.linspace(-40.0, 40.0, 9)
    ant = np.stack([np.full(9, -4000.0), y, np.full(9, 3000.0)], 1)
    b = af.pulse_bins(ant, 9.5e9, 1e6, 200, 128, 0.5, e1=(0.0, 1.0, 0.0))
    fc = 9.5e9 + 100 * 1e6
    assert np.allclose(b, -2 * fc * y / (C * np.hypot(5000.0, y)) * 0.5 * 128)
    assert b[4] == 0.0


def test_deramp_phase_centre_and_plane():
    """psi vanishes at the scene centre, and its gradient there is zero (the planar term cancels the spherical
    one to first order)."""
    ant = np.stack([np.full(5, -4000.0), np.linspace(-20, 20, 5), np.full(5, 3000.0)], 1)
    psi = af.deramp_phase(ant, 9.5e9, 1e6, 200, 64, 64, 0.5, 0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    assert psi.shape == (64, 64)
    assert abs(psi[32, 32]) < 1e-9
    assert abs(psi[33, 32] - psi[31, 32]) < 1e-3 and abs(psi[32, 33] - psi[32, 31]) < 1e-3
    assert abs(psi[0, 0]) > 1e-2                      # curvature away from the centre
