"""Synthetic workload shapes (CPU): the uneven-size variant of the
full-hemisphere workload populates every iteration-tile bucket with a long
tail of sizes, as the batched-buckets configuration describes (50-500)."""

import numpy as np

from hyperalignment_amd import synth


def test_tail_radii_sizes():
    # order 5 with the radius cap doubled with the vertex spacing (46 mm for the
    # 23 mm of the order-6 workload): the same size distribution, 4x faster
    coords, s = synth.searchlights(5, 46.0, tail_seed=2024)
    k = s.sizes
    assert k.min() >= 30 and k.max() <= 560
    assert 70 <= np.median(k) <= 110          # most searchlights are small
    assert np.percentile(k, 90) >= 250        # and a long tail of large ones
    kp = (k + 31) // 32 * 32
    counts = np.bincount(np.searchsorted([64, 128, 256, 384, 512], kp, side="left"), minlength=6)
    assert np.all(counts[:5] > 0), counts
    # every searchlight holds its centre, and lists are sorted and distinct
    for c in range(0, s.count, 997):
        lst = s.indices[s.offsets[c]:s.offsets[c + 1]]
        assert c in lst and np.all(np.diff(lst) > 0)
    assert np.all(s.dists < s.radius)


def test_tail_radii_deterministic():
    a = synth.tail_radii(1000, 100.0, 7)
    b = synth.tail_radii(1000, 100.0, 7)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, synth.tail_radii(1000, 100.0, 8))
