"""The KNN drop-ins called with the reference's argument lists
(spinifel/sequential/nearest_neighbor.py:75-145):

* ``calc_eudist_gpu(model_slices, slices, deviceId)`` -> flat squared
  distances, row = image (cudaEuclideanDistance, orientation_matching_sp.cu:612);
* ``calc_argmin_gpu(euDist, n_images, n_refs, n_pixels, deviceId)`` -> the
  nearest reference per image (cudaHeapSort with totalNN = 1, :682);
* ``nearest_neighbor(model_slices, slices, batch_size)`` -> argmin over
  orientations per image.

Inputs restate tests/test_OM.py:51-76 of the reference (uniform random
model slices and images, 128 x 128 = 16,384 pixels, 512 models, 1,024
images) with a seed; the check there, argmin equal to the FP64 argmin of
scikit-learn's squared distances, is kept, and the distances are compared
with scikit-learn's to the FP32 bound.
"""
import numpy as np
import pytest
from sklearn.metrics.pairwise import euclidean_distances

from oracle import matching

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(76)
    model_slices = rng.random((512, 128 * 128))
    slices = rng.random((1024, 128 * 128))
    return model_slices, slices


def test_calc_eudist_and_argmin_gpu_reference_arguments(data):
    from spinifel_amd import nearest_neighbor as nn

    model_slices, slices = data
    deviceId = 0
    print("knn: calc_eudist_gpu 512 x 1024 x 16384", flush=True)
    euDist = nn.calc_eudist_gpu(model_slices, slices, deviceId)
    assert euDist.shape == (slices.shape[0] * model_slices.shape[0],)
    cpu = euclidean_distances(model_slices, slices, squared=True)       # (N_o, N_i)
    gpu = euDist.reshape(slices.shape[0], model_slices.shape[0]).T
    # FP32 output of distances near 2,700: a few float32 ulps
    np.testing.assert_allclose(gpu, cpu, rtol=2e-6)
    # the smallest gap between an image's two nearest references is 0.012
    # here, far above that error, so the argmin must be the FP64 one
    index = nn.calc_argmin_gpu(euDist, slices.shape[0], model_slices.shape[0], slices.shape[1],
                               deviceId)
    assert index.shape == (slices.shape[0],)
    np.testing.assert_array_equal(index, np.argmin(cpu, axis=0))
    # the same FP32 matrix: the first index of its minimum per image
    np.testing.assert_array_equal(index, np.argmin(gpu, axis=0))


def test_calc_argmin_gpu_ties_first_index():
    from spinifel_amd import nearest_neighbor as nn

    # every reference at the same distance: the lowest index wins (.cu:549-557)
    eu = np.full((3, 40), 5.0, np.float32)
    eu[1, 7] = eu[1, 31] = 1.0
    idx = nn.calc_argmin_gpu(eu.ravel(), 3, 40, 16, 0)
    np.testing.assert_array_equal(idx, [0, 7, 0])


def test_nearest_neighbor_reference_arguments(data):
    from spinifel_amd import nearest_neighbor as nn

    model_slices, slices = data
    print("knn: nearest_neighbor 512 x 1024 x 16384", flush=True)
    index = nn.nearest_neighbor(model_slices, slices, 128)
    np.testing.assert_array_equal(index, matching.nearest_neighbor(model_slices, slices))


@pytest.mark.parametrize("totalNN", [1, 3, 8, 9, 100, 2500, 3000])
def test_heap_sort_any_total_nn(totalNN):
    """mtip_heap_sort for any totalNN (cudaHeapSort, orientation_matching_sp.cu:682,
    takes any count): the totalNN smallest per row, ascending, lowest index on
    ties (entries repeat values on purpose), -1 past the row's length."""
    import torch
    from spinifel_amd import engine
    rng = np.random.default_rng(totalNN)
    n_ref = 2600
    eu = rng.integers(0, 500, (37, n_ref)).astype(np.float32)    # many ties
    eu[5, :] = 7.0                                                # a row of equal values
    got = engine.heap_sort(torch.as_tensor(eu).cuda(), totalNN).cpu().numpy()
    want = np.argsort(eu, axis=1, kind="stable")[:, :totalNN]
    if totalNN > n_ref:
        want = np.concatenate([want, np.full((37, totalNN - n_ref), -1)], axis=1)
    np.testing.assert_array_equal(got, want)


def test_euclidean_distance_rows_not_a_multiple_of_32():
    """mtip_euclidean_distance on rows of 36 and 100 pixels (multiples of 4,
    not of 32) against numpy FP64: the part of the last 128-byte step past the
    row reads as zero."""
    import torch
    from spinifel_amd import engine
    rng = np.random.default_rng(3)
    for n_pix in (36, 100):
        ref = rng.random((300, n_pix)).astype(np.float32)
        dat = rng.random((170, n_pix)).astype(np.float32)
        got = engine.euclidean_distance(torch.from_numpy(ref).cuda(),
                                        torch.from_numpy(dat).cuda()).cpu().numpy()
        want = ((dat.astype(np.float64)[:, None, :] - ref.astype(np.float64)[None]) ** 2).sum(2)
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5)
