"""ORACLE -- test infrastructure, not product code.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module.

numpy restatement of spinifel's data preparation (spinifel/prep.py:36-134):
``bin2x2_sum`` / ``bin2x2_mean`` / ``bin2x2_index`` and their 2^n repeats,
``clipping`` / ``clipping_index``, the ``binning_*`` compositions, and
``compute_pixel_distance`` (:170-172).  The HDF5 loaders (:142-187) are
not restated: h5py is absent here.

Parity status: pinned by brute force (explicit loops over each output
pixel's block, in the reference's summation order), by conservation (the
binned sum of a clipped array keeps its total), constant inputs, and the
clipping windows' positions on index ramps (tests/test_oracle_prep.py).
"""
import numpy as np


def bin2x2_sum(arr):
    """prep.py:36-40."""
    return arr[..., ::2, ::2] + arr[..., 1::2, ::2] + arr[..., ::2, 1::2] + arr[..., 1::2, 1::2]


def bin2x2_mean(arr):
    """prep.py:44-45."""
    return bin2x2_sum(arr) / 4


def bin2x2_index(arr):
    """prep.py:49-53."""
    arr = np.minimum(arr[..., ::2, :], arr[..., 1::2, :])
    arr = np.minimum(arr[..., ::2], arr[..., 1::2])
    return arr // 2


def bin2nx2n(arr, n, mode):
    f = {"sum": bin2x2_sum, "mean": bin2x2_mean, "index": bin2x2_index}[mode]
    for _ in range(n):
        arr = f(arr)
    return arr


def clipping(arr, n):
    """prep.py:73-105 (the 1- and 4-panel windows; zeros otherwise)."""
    n = 2 ** n
    sa, sb = arr.shape[-2:]
    narr = np.zeros(arr.shape[:-2] + (sa // n, sb // n), dtype=arr.dtype)
    if narr.shape[1] == 4:
        narr[..., 0, :, :] = arr[..., 0, -sa // n:, -sb // n:]
        narr[..., 1, :, :] = arr[..., 1, -sa // n:, :sb // n]
        narr[..., 2, :, :] = arr[..., 2, -sa // n:, -sb // n:]
        narr[..., 3, :, :] = arr[..., 3, -sa // n:, :sb // n]
    elif narr.shape[1] == 1:
        n *= 2
        narr[..., 0, :, :] = arr[..., 0, sa // 2 - sa // n:sa // 2 + sa // n,
                                 sb // 2 - sb // n:sb // 2 + sb // n]
    return narr


def clipping_index(arr, n):
    """prep.py:108-113."""
    arr = clipping(arr, n)
    arr -= arr.min()
    return arr


def binning(arr, n_clip, n_bin, mode):
    """binning_sum / binning_mean / binning_index (prep.py:116-134)."""
    if mode == "index":
        return bin2nx2n(clipping_index(arr, n_clip), n_bin, "index")
    return bin2nx2n(clipping(arr, n_clip), n_bin, mode)


def compute_pixel_distance(pixel_position_reciprocal):
    """prep.py:170-172."""
    return np.sqrt((pixel_position_reciprocal ** 2).sum(axis=0))
