"""Pins the oracle's class path (SNM.slicing_and_match with NUFFT.forward):
one batch equals the function path; a repeated grid gives both copies the
same scale, so every argmin lands in the first copy at the same index."""
import numpy as np

from oracle import matching
from spinifel_amd import synth


def _case(n_ref=60, n_img=40, seed=4):
    ac = synth.make_model(17, 20, seed)
    ppr, pdr = synth.detector(1, 8, 8)
    ref = synth.random_quaternions(n_ref, np.random.default_rng(seed + 1))
    true = synth.random_quaternions(n_img, np.random.default_rng(seed + 2))
    sl = matching.model_slices(ac, true, ppr, pdr.max(), 1, None, 1e-9)
    imgs = synth.poisson_images(sl, 3.0, np.random.default_rng(seed + 3)).reshape(n_img, 1, 8, 8)
    return ac, ppr, pdr, ref, imgs


def test_one_batch_equals_function_path():
    ac, ppr, pdr, ref, imgs = _case()
    a = matching.snm_slicing_and_match(ac, imgs, ppr, pdr, ref, len(ref), eps=1e-9,
                                       return_details=True)
    b = matching.slicing_and_match(ac, imgs, ppr, pdr, ref, eps=1e-9, return_details=True)
    np.testing.assert_array_equal(a["index"], b["index"])
    np.testing.assert_allclose(a["mult"], b["alpha"], rtol=1e-12)


def test_repeated_grid_same_scale_first_copy_wins():
    ac, ppr, pdr, ref, imgs = _case()
    one = matching.snm_slicing_and_match(ac, imgs, ppr, pdr, ref, len(ref), eps=1e-9,
                                         return_details=True)
    two = matching.snm_slicing_and_match(ac, imgs, ppr, pdr, np.concatenate([ref, ref]),
                                         len(ref), eps=1e-9, return_details=True)
    np.testing.assert_allclose(two["mult"][:len(ref)], two["mult"][len(ref):], rtol=1e-12)
    np.testing.assert_array_equal(two["index"], one["index"])


def test_batches_scale_to_the_data_std():
    ac, ppr, pdr, ref, imgs = _case()
    r = matching.snm_slicing_and_match(ac, imgs, ppr, pdr, ref, 20, eps=1e-9,
                                       return_details=True)
    data_std = imgs.astype(np.float64).std()
    for b in range(3):
        np.testing.assert_allclose(r["model_slices"][20 * b:20 * (b + 1)].std(), data_std,
                                   rtol=1e-12)
    assert matching.snm_slicing_and_match(ac, imgs[:0], ppr, pdr, ref, 20).shape == (0, 4)
