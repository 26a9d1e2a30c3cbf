"""Distances and argmin, drop-in for spinifel/sequential/nearest_neighbor.py.

``nearest_neighbor(model_slices, slices, batch_size)`` keeps the reference
signature and result (spinifel/sequential/nearest_neighbor.py:103-145): the
index of the closest model slice for every image, first index on ties
(``np.argmin(axis=0)``, :135).  Here the cross term runs on MFMA with a
fused top-K and the candidates are rescored in FP64, so the full
``(N_o, N_i)`` matrix is never materialised.  ``batch_size`` is accepted
for compatibility; the kernels tile internally.

``calc_eudist_gpu`` / ``calc_argmin_gpu`` keep the signatures of the
reference's KNN pair (:75-98) over ``mtip_euclidean_distance`` /
``mtip_heap_sort``.
"""
import numpy as np
import torch

from . import engine
from .settings import settings as _settings


def _dev(deviceId=None):
    engine.require_gpu()
    return torch.device("cuda", _settings.device if deviceId is None else deviceId)


def _rows_f32(a, dev, N_p_pad):
    a = np.ascontiguousarray(a, dtype=np.float32)
    t = torch.zeros((a.shape[0], N_p_pad), dtype=torch.float32, device=dev)
    t[:, :a.shape[1]] = torch.from_numpy(a).to(dev)
    return t


def calc_eudist_gpu(model_slices, slices, deviceId=None):
    """Squared distances, flat (N_images * N_model,) float32, row = image
    (cudaEuclideanDistance, orientation_matching_sp.cu:612).

    Both sets are first centred on the mean model slice: the distance is
    invariant to a common shift, and without the shift the FP32 cross term
    2 x.y nearly cancels ||x||^2 + ||y||^2 (uniform random images in [0, 1)
    over 16,384 pixels: about 8,200 against a distance of 2,700), so its
    rounding would reach a few 1e-5 of the distance instead of about 1e-7."""
    dev = _dev(deviceId)
    model_slices = np.asarray(model_slices).reshape(model_slices.shape[0], -1)
    slices = np.asarray(slices).reshape(slices.shape[0], -1)
    n_pix = model_slices.shape[1]
    P = engine.padded(n_pix, 32)
    ref = _rows_f32(model_slices, dev, P)
    data = _rows_f32(slices, dev, P)
    c = engine.col_mean(ref, n_pix)
    ref = engine.residual(ref, "fp32", n_pix, center=c).data
    data = engine.residual(data, "fp32", n_pix, center=c).data
    return engine.euclidean_distance(ref, data).cpu().numpy().reshape(-1)


def calc_argmin_gpu(euDist, n_images, n_refs, n_pixels, deviceId=None, totalNN=1):
    """Indices of the totalNN nearest references per image (cudaHeapSort, .cu:682)."""
    dev = _dev(deviceId)
    d = torch.as_tensor(np.ascontiguousarray(euDist, np.float32).reshape(n_images, n_refs)).to(dev)
    idx = engine.heap_sort(d, totalNN).cpu().numpy()
    return idx.reshape(-1) if totalNN == 1 else idx


def nearest_neighbor(model_slices, slices, batch_size=None, precision="fp32", topk=8):
    """argmin over model slices for every image (int64 (N_images,)).

    The model slices are taken as given (already scaled).  Candidates come
    from the MFMA pass in ``precision`` and are rescored in FP64.
    """
    dev = _dev()
    model_slices = np.asarray(model_slices).reshape(model_slices.shape[0], -1)
    slices = np.asarray(slices).reshape(slices.shape[0], -1)
    imgs = engine.ImageSet(slices, dev)

    S = engine.SliceSet.from_rows(_rows_f32(model_slices, dev, imgs.ld), imgs.N_p)
    one = torch.ones(1, dtype=torch.float64, device=dev)
    K = max(1, min(8, int(topk)))
    K = 1 << (K - 1).bit_length()
    r = engine.match_resident(S, imgs, precision=precision, K=K, rescore_fp64=True, alpha=one)
    return r["index"].long().cpu().numpy()
