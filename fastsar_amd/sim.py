"""Synthetic spotlight collections: geometry, point scenes and frequency-domain phase history.

Conventions (fastsar/sim.py:3-5, 17-105): sample k of pulse p is at frequency f_k = fmin + k df; the history is
motion compensated to the scene reference point (the origin), so a scatterer at x with complex amplitude a adds
a exp(-j 4 pi f_k (|x - a_p| - |a_p|) / c).

  Collect            fmin, df, K, ant [P, 3], res                         (fastsar/sim.py:17-38)
  make_collect       broadside linear-track collect of the reference's tests (fastsar/sim.py:46-67)
  delta_range        |x - a_p| - |a_p| in float64                         (fastsar/sim.py:70-73)
  simulate_brute     the direct float64 sum, in blocks of scatterers      (fastsar/sim.py:90-100)
  add_noise          i.i.d. complex Gaussian samples                      (fastsar/sim.py:103-105)
  spotlight          the benchmark geometries: a linear track at slant range R and grazing angle psi spanning an
                     aperture angle, with evenly spaced pulses, fc = 9.6 GHz and 300 MHz of band by default
  point_scene        seeded uniform point targets with complex Gaussian amplitudes
"""
from dataclasses import dataclass

import numpy as np

C = 299792458.0


@dataclass
class Collect:
    fmin: float          # Hz, frequency of sample k = 0
    df: float            # Hz, frequency step
    K: int               # frequency samples per pulse
    ant: np.ndarray      # [P, 3] antenna phase centres, metres, scene-centred
    res: float           # nominal resolution, metres

    @property
    def Np(self):
        return self.ant.shape[0]

    @property
    def fref(self):
        """Frequency of sample K // 2, the carrier after range compression."""
        return self.fmin + (self.K // 2) * self.df

    @property
    def freqs(self):
        return self.fmin + self.df * np.arange(self.K)


def _even(n):
    n = int(np.ceil(n))
    return n + (n % 2)


def make_collect(res=0.3, scene=150.0, r0=10e3, graze_deg=30.0, fc=9.6e9, margin=1.25, offset=(0.0, 0.0, 0.0)):
    """Broadside linear track along +y looking along +x, sampled so that a square patch of side `scene` (times a
    margin of 1.25 on its diagonal) is unaliased in range and along track. (fastsar/sim.py:46-67)"""
    lam = C / fc
    extent = margin * np.sqrt(2.0) * scene
    df = C / (2.0 * extent)
    K = _even((C / (2.0 * res)) / df)
    psi = np.deg2rad(graze_deg)
    length = r0 * lam / (2.0 * res)
    dy = lam * r0 / (2.0 * extent)
    P = _even(length / dy)
    y = (np.arange(P) - (P - 1) / 2.0) * dy
    ant = np.stack([np.full(P, -r0 * np.cos(psi)), y, np.full(P, r0 * np.sin(psi))], axis=1)
    ant = ant + np.asarray(offset, np.float64)[None, :]
    return Collect(fmin=fc - (K // 2) * df, df=df, K=K, ant=ant, res=res)


def spotlight(P, K, aperture_deg, r0=10e3, graze_deg=30.0, fc=9.6e9, bandwidth=300e6, squint_deg=0.0):
    """Linear track at slant range r0 from the scene centre and grazing angle psi: a_p = (-r0 cos psi, y_p, r0 sin
    psi), y_p evenly spaced over r0 cos psi tan(+-aperture / 2) (the aperture angle as seen from the scene centre in
    the ground plane), shifted along track by r0 cos psi tan(squint). df = bandwidth / K, fmin = fc - (K // 2) df."""
    psi = np.deg2rad(graze_deg)
    half = r0 * np.cos(psi) * np.tan(np.deg2rad(aperture_deg) / 2.0)
    y = np.linspace(-half, half, P) + r0 * np.cos(psi) * np.tan(np.deg2rad(squint_deg))
    ant = np.stack([np.full(P, -r0 * np.cos(psi)), y, np.full(P, r0 * np.sin(psi))], axis=1)
    df = bandwidth / K
    return Collect(fmin=fc - (K // 2) * df, df=df, K=K, ant=ant, res=C / (2.0 * bandwidth))


def point_scene(rng, n, ext1, ext2, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0)):
    """n targets uniform over [-ext1/2, ext1/2] along e1 and [-ext2/2, ext2/2] along e2 (z = 0 plane), with complex
    Gaussian amplitudes (Rayleigh magnitude, uniform phase): (pos [n, 3], amp [n] complex128)."""
    e1, e2 = np.asarray(e1, np.float64), np.asarray(e2, np.float64)
    s1 = rng.uniform(-ext1 / 2, ext1 / 2, n)
    s2 = rng.uniform(-ext2 / 2, ext2 / 2, n)
    pos = s1[:, None] * e1 + s2[:, None] * e2
    amp = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2.0)
    return pos, amp


def delta_range(ant, pos):
    """dR[p, n] = |pos_n - ant_p| - |ant_p| in float64. (fastsar/sim.py:70-73)"""
    d = pos[None, :, :] - ant[:, None, :]
    return np.sqrt(np.einsum('pnk,pnk->pn', d, d)) - np.linalg.norm(ant, axis=1)[:, None]


def simulate_brute(col, pos, amp):
    """The phase history [P, K] complex128 by the direct float64 sum over scatterers, in blocks that keep the
    temporaries near 64 MB. (fastsar/sim.py:90-100)"""
    dr = delta_range(col.ant, np.asarray(pos, np.float64))
    amp = np.asarray(amp)
    out = np.zeros((col.Np, col.K), np.complex128)
    nb = max(1, (1 << 22) // (col.Np * col.K))
    f = col.freqs
    for n0 in range(0, dr.shape[1], nb):
        ph = (-4.0 * np.pi / C) * f[None, :, None] * dr[:, None, n0:n0 + nb]
        out += (amp[None, None, n0:n0 + nb] * np.exp(1j * ph)).sum(-1)
    return out


def add_noise(S, sigma, rng):
    """S plus complex Gaussian noise of standard deviation sigma per complex sample. (fastsar/sim.py:103-105)"""
    n = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    return S + (sigma / np.sqrt(2.0)) * n


def clutter_sigma(amp, db):
    """Standard deviation of white clutter whose power per sample is `db` below the targets' summed power per
    sample (sum |a|^2, the mean of |sum a exp(...)|^2 over random phases)."""
    return float(np.sqrt(np.sum(np.abs(amp) ** 2) * 10.0 ** (db / 10.0)))


def config1(seed=1, P=512, K=1024, targets=25, clutter_db=-30.0, side=256.0):
    """The small spotlight scene: X band, 300 MHz, P pulses x K samples over a 3 degree aperture at 10 km slant range
    and 30 degrees grazing; `targets` points uniform in a side x side square with Rayleigh amplitudes plus white
    clutter clutter_db below the targets' power. -> (col, S complex64 [P, K], pos, amp, grid dict)."""
    rng = np.random.default_rng(seed)
    col = spotlight(P, K, 3.0)
    pos, amp = point_scene(rng, targets, side, side)
    S = simulate_brute(col, pos, amp)
    if clutter_db is not None:
        S = add_noise(S, clutter_sigma(amp, clutter_db), rng)
    grid = dict(nx=512, ny=512, spx=0.5, spy=0.5, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
    return col, S.astype(np.complex64), pos, amp, grid


def uneven_batch(seed=5, n=48, sides=(256, 4096), pulses=(512, 8192), samples=(1024, 4096), sp=0.5, r0=10e3,
                 graze_deg=30.0, fc=9.6e9, squint_deg=2.0, targets=20, clutter_db=-30.0, margin=1.25,
                 max_bandwidth=None, max_aperture_deg=6.0):
    """n images of uneven size for the batched workload: square sides, pulse counts and sample counts drawn
    log-uniformly (seeded); even-numbered images spotlight, odd ones stripmap-like patches seen at squint_deg. Each
    collection is sampled for its image: frequency spacing and along-track spacing leave the image's extent (times
    margin on the diagonal) unambiguous, finer where K or P would otherwise span more than max_bandwidth (default
    c / (2 sp): a range resolution no finer than the pixel, as the 512^2 and 12288 x 8192 scenes have) or
    max_aperture_deg, so the bandwidth and the aperture follow from K and P. -> list of dict(col, grid, pos, amp,
    sigma, squint)."""
    rng = np.random.default_rng(seed)
    lam = C / fc
    psi = np.deg2rad(graze_deg)
    max_bandwidth = C / (2.0 * sp) if max_bandwidth is None else max_bandwidth
    out = []
    for i in range(n):
        side = _even(2.0 ** rng.uniform(np.log2(sides[0]), np.log2(sides[1])))
        P = _even(2.0 ** rng.uniform(np.log2(pulses[0]), np.log2(pulses[1])))
        K = _even(2.0 ** rng.uniform(np.log2(samples[0]), np.log2(samples[1])))
        extent = margin * np.sqrt(2.0) * side * sp
        df = min(C / (2.0 * extent), max_bandwidth / K)
        dy = min(lam * r0 / (2.0 * extent), 2.0 * r0 * np.cos(psi) * np.tan(np.deg2rad(max_aperture_deg) / 2.0) / (P - 1))
        half = 0.5 * dy * (P - 1)
        squint = 0.0 if i % 2 == 0 else squint_deg
        ap = np.rad2deg(2.0 * np.arctan(half / (r0 * np.cos(psi))))
        col = spotlight(P, K, ap, r0=r0, graze_deg=graze_deg, fc=fc, bandwidth=K * df, squint_deg=squint)
        grid = dict(nx=side, ny=side, spx=sp, spy=sp, e1=(0.0, 1.0, 0.0), e2=(1.0, 0.0, 0.0))
        pos, amp = point_scene(rng, targets, 0.8 * side * sp, 0.8 * side * sp)
        out.append(dict(col=col, grid=grid, pos=pos, amp=amp, sigma=clutter_sigma(amp, clutter_db), squint=squint))
    return out
