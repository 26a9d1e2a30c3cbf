"""Pins the oracle's distances / argmin / scaling and its slicing_and_match."""
import numpy as np
import pytest
from sklearn.metrics.pairwise import euclidean_distances as sk_euclidean

from oracle import geometry, matching
from spinifel_amd import synth


def test_distances_equal_sklearn():
    rng = np.random.default_rng(0)
    X, Y = rng.random((40, 300)), rng.random((25, 300))
    np.testing.assert_allclose(matching.euclidean_distances(X, Y), sk_euclidean(X, Y), rtol=1e-12)
    np.testing.assert_allclose(matching.euclidean_distances(X, Y, squared=True),
                               sk_euclidean(X, Y, squared=True), rtol=1e-12)
    brute = ((X[:, None, :] - Y[None, :, :]) ** 2).sum(-1)
    np.testing.assert_allclose(matching.euclidean_distances(X, Y, squared=True), brute, rtol=1e-10)


def test_argmin_planted_and_first_occurrence():
    rng = np.random.default_rng(1)
    S = rng.random((30, 64))
    pick = rng.integers(0, 30, 20)
    assert (matching.nearest_neighbor(S, S[pick]) == pick).all()
    S2 = np.concatenate([S, S])       # exact duplicates: the lower index wins
    assert (matching.nearest_neighbor(S2, S[pick]) == pick).all()


def test_topk_squared_brute_force():
    rng = np.random.default_rng(2)
    S, D = rng.random((50, 40)), rng.random((12, 40))
    idx, dd = matching.topk_squared(S, D, 5)
    brute = ((S[:, None, :] - D[None, :, :]) ** 2).sum(-1)
    order = np.argsort(brute, axis=0, kind="stable")[:5].T
    np.testing.assert_array_equal(idx, order)
    np.testing.assert_allclose(dd, np.take_along_axis(brute, order.T, 0).T, rtol=1e-10)


def test_scaling_ratio_is_population_std():
    rng = np.random.default_rng(3)
    a, b = rng.random((5, 7)), rng.random((9, 7)) * 3
    assert matching.scaling_ratio(a, b) == pytest.approx(np.std(a) / np.std(b))
    assert matching.scaling_ratio(a, b) == pytest.approx(
        np.sqrt(((a - a.mean()) ** 2).mean()) / np.sqrt(((b - b.mean()) ** 2).mean()))


def test_slicing_and_match_planted_images():
    """Noise-free images equal to scaled model slices at reference j match j."""
    ac = synth.make_model(17, 20, 4)
    ppr, pdr = synth.detector(1, 8, 8)
    E = pdr.max()
    ref = synth.random_quaternions(48, np.random.default_rng(5))
    S = matching.model_slices(ac, ref, ppr, E, 1, None, 1e-9)
    pick = np.random.default_rng(6).permutation(48)   # all slices: std ratio is exactly 3
    imgs = (3.0 * S[pick]).reshape((48, 1, 8, 8))
    r = matching.slicing_and_match(ac, imgs, ppr, pdr, ref, return_details=True)
    assert (r["index"] == pick).all()
    assert r["alpha"] == pytest.approx(3.0, rel=1e-9)   # planted at eps 1e-9, matched at 1e-12
    np.testing.assert_allclose(r["min_dist"], 0.0, atol=1e-6 * np.abs(imgs).max())
    np.testing.assert_array_equal(matching.slicing_and_match(ac, imgs, ppr, pdr, ref), ref[pick])
    assert matching.slicing_and_match(ac, imgs[:0], ppr, pdr, ref).shape == (0, 4)


def test_gen_model_slices_batch_consistency():
    ac = synth.make_model(15, 15, 7)
    ppr, pdr = synth.detector(1, 4, 8)
    E = pdr.max()
    ref = synth.random_quaternions(6, np.random.default_rng(8))
    a = matching.gen_model_slices(ac, ref, ppr, E, 1, None, 32)
    b = matching.gen_model_slices_batch(ac, ref, ppr, E, 1, None, 32, batch_size=3)
    np.testing.assert_allclose(a, b, rtol=1e-12)
    assert a.shape == (6 * 32,)
    with pytest.raises(AssertionError):
        matching.gen_model_slices_batch(ac, ref, ppr, E, 1, None, 32, batch_size=4)
    H, K, L = geometry.gen_nonuniform_normalized_positions(ref, ppr, E, 1)
    from oracle import es_nufft
    np.testing.assert_allclose(a, es_nufft.nudft3d2(H, K, L, ac.astype(np.float64)).real,
                               rtol=1e-9, atol=1e-9 * np.abs(a).max())
