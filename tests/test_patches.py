"""Host side of the patch mosaic (fastsar_amd.patches) on the CPU: dropped pulses put back on an even spacing, the
pulses a beam serves a patch with, and the levels a patch takes (fastsar/patches.py:162-171, 551-590)."""
import numpy as np

from fastsar_amd import patches as pt


def test_fill_gaps_puts_dropped_pulses_back():
    rng = np.random.default_rng(1)
    full = np.stack([np.full(300, -5e3), np.arange(300) * 0.37, np.full(300, 2e3)], 1) + rng.normal(0, 1e-4, (300, 3))
    keep = np.ones(300, bool)
    keep[[40, 41, 42, 100, 200, 201]] = False
    pos, at = pt.fill_gaps(full[keep])
    assert len(pos) == 300 and np.array_equal(at, np.flatnonzero(keep))
    np.testing.assert_allclose(pos[at], full[keep])
    steps = np.linalg.norm(np.diff(pos, axis=0), axis=1)
    assert steps.max() / steps.min() < 1.01


def test_fill_gaps_leaves_even_tracks():
    a = np.stack([np.zeros(50), np.linspace(0, 10, 50), np.zeros(50)], 1)
    pos, at = pt.fill_gaps(a)
    assert pos is not None and np.array_equal(pos, a) and np.array_equal(at, np.arange(50))
    still = np.zeros((5, 3))
    assert np.array_equal(pt.fill_gaps(still)[1], np.arange(5))


def test_beam_span_against_every_pulse():
    """the coarse-then-fine search finds the same run as testing every pulse"""
    P = 1000
    ant = np.stack([np.full(P, -4e3), np.linspace(-600, 600, P), np.full(P, 3e3)], 1)

    def beam(idx, pts):
        w = np.asarray(pts)[None] - ant[idx][:, None]
        return 20.0 * w[..., 1] / np.linalg.norm(w, axis=-1)
    for y in (-580.0, -100.0, 0.0, 250.0, 590.0):
        pts = np.array([[0.0, y - 5, 0.0], [0.0, y + 5, 0.0]])
        every = np.flatnonzero((np.abs(beam(np.arange(P), pts)) <= 1.0).any(1))
        got = pt.beam_span(beam, P, pts, 1.0)
        assert got == ((int(every[0]), int(every[-1]) + 1) if every.size else None), (y, got)
    far = np.array([[0.0, 9e3, 0.0]])
    assert pt.beam_span(beam, P, far, 1.0) is None


def test_patch_levels():
    assert pt.patch_levels(64, 16) == 2 and pt.patch_levels(128, 16) == 3 and pt.patch_levels(16, 16) == 1
    assert pt.patch_levels(96, 32) == 1 and pt.patch_levels(96, 32, longest=832) == 2
