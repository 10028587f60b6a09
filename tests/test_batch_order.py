"""How a batch deals its images to streams (fastsar_amd.batch.stream_order), on the host: every image once, largest
first, each to the stream with the least work so far; against round-robin on uneven costs."""
import numpy as np

from fastsar_amd.batch import stream_order


def _loads(cost, order, stream, ns):
    out = [0.0] * ns
    for p, i in enumerate(order):
        out[stream[p]] += cost[i]
    return out


def test_every_image_once_largest_first():
    cost = [3.0, 9.0, 1.0, 9.0, 4.0]
    order, stream = stream_order(cost, 2)
    assert order == [1, 3, 4, 0, 2]               # ties kept in index order
    assert stream == [0, 1, 0, 1, 1]              # 9 | 9, then 4 -> 0 (tie, lower), 3 -> 1, 1 -> 1 (12 against 13)
    assert sorted(order) == list(range(len(cost)))


def test_balanced_against_round_robin():
    rng = np.random.default_rng(5)
    for ns in (2, 3, 4, 8):
        cost = list(rng.lognormal(0.0, 1.0, 48))
        order, stream = stream_order(cost, ns)
        got = _loads(cost, order, stream, ns)
        rr = _loads(cost, list(range(48)), [i % ns for i in range(48)], ns)
        # greedy largest-first: the fullest stream exceeds the mean by at most the largest single cost
        assert max(got) <= sum(cost) / ns + max(cost) + 1e-9
        assert max(got) - min(got) <= max(rr) - min(rr) + 1e-9


def test_more_streams_than_images():
    order, stream = stream_order([1.0, 2.0], 4)
    assert order == [1, 0] and stream == [0, 1]
