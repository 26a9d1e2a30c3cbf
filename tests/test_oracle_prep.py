"""Pins the oracle's data preparation (binning / clipping) by brute force,
conservation and closed forms."""
import numpy as np
import pytest

from oracle import prep


def _brute_sum(arr, n):
    """Each output pixel as its block's sum in the reference's order, by loops."""
    def blk(a, r, c, L):
        if L == 0:
            return a[r, c]
        h = 1 << (L - 1)
        return ((blk(a, r, c, L - 1) + blk(a, r + h, c, L - 1)) + blk(a, r, c + h, L - 1)) \
            + blk(a, r + h, c + h, L - 1)
    lead = arr.shape[:-2]
    sa, sb = arr.shape[-2:]
    f = 1 << n
    out = np.zeros(lead + (sa // f, sb // f), dtype=arr.dtype)
    for idx in np.ndindex(*lead):
        for r in range(sa // f):
            for c in range(sb // f):
                out[idx + (r, c)] = blk(arr[idx], r * f, c * f, n)
    return out


@pytest.mark.parametrize("n", [1, 2])
def test_sum_matches_brute_force_bitwise(n):
    a = np.random.default_rng(n).standard_normal((2, 1, 8, 12)).astype(np.float32)
    got = prep.bin2nx2n(a, n, "sum")
    np.testing.assert_array_equal(got, _brute_sum(a, n))


def test_mean_and_index_closed_forms():
    a = np.full((3, 4, 8, 8), 2.5)
    np.testing.assert_array_equal(prep.bin2nx2n(a, 2, "mean"), np.full((3, 4, 2, 2), 2.5))
    # index ramps: the min of each block, halved per level
    r = np.arange(16)
    idx = np.stack(np.meshgrid(r, r, indexing="ij"))[:, None]        # (2, 1, 16, 16)
    out = prep.bin2nx2n(idx, 2, "index")
    np.testing.assert_array_equal(out[0, 0], np.repeat(np.arange(4)[:, None], 4, 1))
    np.testing.assert_array_equal(out[1, 0], np.repeat(np.arange(4)[None, :], 4, 0))
    # floor division of negatives
    np.testing.assert_array_equal(prep.bin2x2_index(np.array([[[[-3, 5], [7, 9]]]])),
                                  np.array([[[[-2]]]]))


def test_clipping_windows_and_conservation():
    rng = np.random.default_rng(3)
    mono = rng.uniform(size=(2, 1, 16, 24))
    c = prep.clipping(mono, 1)
    assert c.shape == (2, 1, 8, 12)
    np.testing.assert_array_equal(c, mono[:, :, 4:12, 6:18])
    four = rng.uniform(size=(2, 4, 8, 8))
    c4 = prep.clipping(four, 1)
    np.testing.assert_array_equal(c4[:, 0], four[:, 0, 4:, 4:])
    np.testing.assert_array_equal(c4[:, 1], four[:, 1, 4:, :4])
    np.testing.assert_array_equal(c4[:, 3], four[:, 3, 4:, :4])
    assert not prep.clipping(rng.uniform(size=(1, 2, 8, 8)), 1).any()
    b = prep.binning(mono, 1, 2, "sum")
    assert b.shape == (2, 1, 2, 3)
    np.testing.assert_allclose(b.sum(axis=(-1, -2)), c.sum(axis=(-1, -2)), rtol=1e-12)
    ci = prep.clipping_index(np.arange(2 * 16 * 16).reshape(2, 1, 16, 16), 1)
    assert ci.min() == 0


def test_pixel_distance():
    p = np.random.default_rng(4).standard_normal((3, 1, 4, 5))
    np.testing.assert_allclose(prep.compute_pixel_distance(p),
                               np.linalg.norm(p, axis=0), rtol=1e-14)
