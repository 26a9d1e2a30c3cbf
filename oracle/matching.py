"""ORACLE -- test infrastructure, not product code.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module.

FP64 CPU restatement of spinifel's orientation-matching path (CPU mode):

* ``slicing_and_match`` follows spinifel/sequential/orientation_matching.py:189-256
  (batched slicing :230-241, global std scaling :245-248, nearest neighbour
  :252, gather :256) -- the body behind ``match``
  (spinifel/mpi/orientation_matching.py:8-16).
* ``euclidean_distances`` restates scikit-learn's dense formula used at
  spinifel/sequential/nearest_neighbor.py:134:
  ``sqrt(max(||x||^2 - 2 x.y + ||y||^2, 0))`` in FP64.
* ``nearest_neighbor`` is the CPU branch, nearest_neighbor.py:131-145
  (``np.argmin(axis=0)``: first index of the minimum).
* ``gen_model_slices`` / ``gen_model_slices_batch`` implement the evident
  intent of spinifel/slicing.py:10-95 (the reference body passes nine
  arguments to the seven-parameter ``forward``; see the docstring there).

Parity status: distances pinned against scikit-learn itself (the
reference's dependency, importable here) and brute force ``sum (s - d)^2``;
argmin pinned by planted images (noise-free image = slice j -> j) and the
first-occurrence tie rule; the scaling by ``np.std`` closed form.  The
default orientation grid ``skopi.get_uniform_quat`` is not restated
(skopi is absent): callers must pass ``ref_orientations``.
"""
import numpy as np

from . import es_nufft, geometry


def forward(ugrid, H_, K_, L_, support, use_recip_sym, N, eps=1e-12):
    """spinifel/autocorrelation.py:32-61 (type-2 NUFFT of the model)."""
    assert np.max(np.abs(np.array([H_, K_, L_]))) < 3 * np.pi
    if use_recip_sym:
        assert np.all(np.isreal(ugrid))
    if support is not None:
        ugrid = ugrid * support
    out = es_nufft.nufft3d2(H_, K_, L_, np.real(ugrid), -1, eps)
    assert out.shape[0] == N
    return out


def euclidean_distances(X, Y, squared=False):
    """scikit-learn's dense euclidean_distances, FP64."""
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    XX = np.einsum("ij,ij->i", X, X)[:, None]
    YY = np.einsum("ij,ij->i", Y, Y)[None, :]
    D = -2.0 * (X @ Y.T)
    D += XX
    D += YY
    np.maximum(D, 0, out=D)
    return D if squared else np.sqrt(D)


def nearest_neighbor(model_slices, slices, chunk=2048, return_dist=False):
    """argmin over orientations (axis 0) per image; first occurrence wins."""
    model_slices = np.asarray(model_slices, dtype=np.float64)
    slices = np.asarray(slices, dtype=np.float64)
    n_img = slices.shape[0]
    index = np.zeros(n_img, dtype=np.int64)
    mind = np.zeros(n_img)
    for s in range(0, n_img, chunk):
        d = euclidean_distances(model_slices, slices[s:s + chunk])
        index[s:s + chunk] = np.argmin(d, axis=0)
        mind[s:s + chunk] = d[index[s:s + chunk], np.arange(d.shape[1])]
    return (index, mind) if return_dist else index


def topk_squared(model_slices, slices, k, chunk=1024):
    """Per image the k smallest squared distances (FP64, direct sum of squares).

    Returns (idx (N_i, k), dist2 (N_i, k)), ascending, lowest index first on ties.
    """
    S = np.asarray(model_slices, dtype=np.float64)
    D = np.asarray(slices, dtype=np.float64)
    n_img = D.shape[0]
    idx = np.zeros((n_img, k), dtype=np.int64)
    dd = np.zeros((n_img, k))
    s2 = np.einsum("ij,ij->i", S, S)
    for s in range(0, n_img, chunk):
        Dc = D[s:s + chunk]
        d2 = s2[:, None] - 2.0 * (S @ Dc.T) + np.einsum("ij,ij->i", Dc, Dc)[None, :]
        order = np.argsort(d2, axis=0, kind="stable")[:k]
        idx[s:s + chunk] = order.T
        dd[s:s + chunk] = np.take_along_axis(d2, order, 0).T
    return idx, dd


def scaling_ratio(slices_, model_slices):
    """``slices_.std() / model_slices.std()`` (population std), FP64."""
    return np.asarray(slices_, np.float64).std() / np.asarray(model_slices, np.float64).std()


def model_slices(ac, ref_orientations, pixel_position_reciprocal, reciprocal_extent,
                 oversampling=1, N_batch_size=None, eps=1e-12):
    """Batched slicing of orientation_matching.py:230-241 -> (N_o, N_p) FP64."""
    ref_orientations = np.asarray(ref_orientations, np.float64).reshape(-1, 4)
    N_o = ref_orientations.shape[0]
    ppr = np.asarray(pixel_position_reciprocal, np.float64)
    N_p = int(np.prod(ppr.shape[1:]))
    if N_batch_size is None:
        N_batch_size = N_o
    assert N_o % N_batch_size == 0, "N_orientations must be divisible by N_batch_size"
    rot = geometry.ref_rotmat(ref_orientations)
    ac = np.asarray(ac)
    assert np.all(np.isreal(ac))
    b, n, w, beta = es_nufft.fine_grid(np.real(ac).astype(np.float64), -1, eps)
    out = np.zeros((N_o, N_p))
    for i in range(N_o // N_batch_size):
        st, en = i * N_batch_size, (i + 1) * N_batch_size
        H, K, L = np.einsum("ijk,k...->ji...", rot[st:en], ppr)
        H_ = H.flatten() / reciprocal_extent * np.pi / oversampling
        K_ = K.flatten() / reciprocal_extent * np.pi / oversampling
        L_ = L.flatten() / reciprocal_extent * np.pi / oversampling
        assert np.max(np.abs(np.array([H_, K_, L_]))) < 3 * np.pi
        out[st:en] = es_nufft.interp(b, H_, K_, L_, w, beta).real.reshape(en - st, N_p)
    return out


def slicing_and_match(ac, slices_, pixel_position_reciprocal, pixel_distance_reciprocal,
                      ref_orientations, N_batch_size=None, oversampling=1, eps=1e-12,
                      return_details=False):
    """Restatement of sequential/orientation_matching.py:189-256.

    Returns ``ref_orientations[index]`` (or a dict of intermediates).
    """
    slices_ = np.asarray(slices_)
    N_slices = slices_.shape[0]
    if not N_slices:
        return np.zeros((0, 4))
    if ref_orientations is None:
        raise ValueError("oracle needs explicit ref_orientations (skopi absent)")
    ref_orientations = np.asarray(ref_orientations)
    reciprocal_extent = np.asarray(pixel_distance_reciprocal).max()
    ms = model_slices(ac, ref_orientations, pixel_position_reciprocal,
                      reciprocal_extent, oversampling, N_batch_size, eps)
    data = slices_.reshape(N_slices, -1)
    alpha = scaling_ratio(data, ms)
    ms *= alpha
    index, mind = nearest_neighbor(ms, data, return_dist=True)
    if return_details:
        return dict(index=index, min_dist=mind, alpha=alpha, model_slices=ms,
                    orientations=ref_orientations[index])
    return ref_orientations[index]


def snm_slicing_and_match(ac, slices_, pixel_position_reciprocal, pixel_distance_reciprocal,
                          ref_orientations, N_batch_size, oversampling=1, eps=1e-12,
                          return_details=False):
    """Restatement of the class path SNM.slicing_and_match
    (sequential/orientation_matching.py:122-186, with NUFFT.forward of
    extern/nufft_ext.py:248-284): per batch of N_batch_size orientations the
    model slices are scaled by ``slices_.std() / batch.std()`` (:165), then
    squared distances ``||x||^2 + ||y||^2 - 2 x.y`` (:99-108) and the argmin
    over all orientations (:179)."""
    slices_ = np.asarray(slices_)
    N_slices = slices_.shape[0]
    if not N_slices:
        return np.zeros((0, 4))
    ref_orientations = np.asarray(ref_orientations)
    N_o = ref_orientations.shape[0]
    assert N_o % N_batch_size == 0, "N_orientations must be divisible by N_batch_size"
    data = slices_.reshape(N_slices, -1).astype(np.float64)
    std = data.std()
    E = np.asarray(pixel_distance_reciprocal).max()
    ms = model_slices(ac, ref_orientations, pixel_position_reciprocal, E, oversampling,
                      N_batch_size, eps)
    mult = np.empty(N_o)
    for b in range(N_o // N_batch_size):
        st, en = b * N_batch_size, (b + 1) * N_batch_size
        mult[st:en] = std / ms[st:en].std()
        ms[st:en] *= mult[st:en, None]
    dist = (ms * ms).sum(1)[:, None] + (data * data).sum(1)[None, :] - 2.0 * ms @ data.T
    index = np.argmin(dist, axis=0)
    if return_details:
        return dict(index=index, mult=mult, model_slices=ms, dist=dist,
                    orientations=ref_orientations[index])
    return ref_orientations[index]


def gen_model_slices(ac, ref_orientations, pixel_position_reciprocal, reciprocal_extent,
                     oversampling, ac_support_size, N_pixels, override_forward_with=None,
                     eps=1e-12):
    """Evident intent of spinifel/slicing.py:10-45: Re(NUFFT) flat (N_o*N_p,)."""
    H_, K_, L_ = geometry.gen_nonuniform_normalized_positions(
        ref_orientations, pixel_position_reciprocal, reciprocal_extent, oversampling)
    N = N_pixels * np.asarray(ref_orientations).reshape(-1, 4).shape[0]
    return forward(np.asarray(ac), H_, K_, L_, 1, True, N, eps).real


def gen_model_slices_batch(ac, ref_orientations, pixel_position_reciprocal,
                           reciprocal_extent, oversampling, ac_support_size, N_pixels,
                           batch_size=None, override_forward_with=None, eps=1e-12):
    """spinifel/slicing.py:50-95."""
    ref_orientations = np.asarray(ref_orientations).reshape(-1, 4)
    N_o = ref_orientations.shape[0]
    bs = N_o if batch_size is None else batch_size
    assert N_o % bs == 0, f"N_orientations ({N_o}) must be divisible by N_batch_size ({bs})"
    out = np.zeros((N_pixels * N_o,))
    for i in range(N_o // bs):
        st, en = i * bs, (i + 1) * bs
        out[st * N_pixels:en * N_pixels] = gen_model_slices(
            ac, ref_orientations[st:en], pixel_position_reciprocal, reciprocal_extent,
            oversampling, ac_support_size, N_pixels, override_forward_with, eps)
    return out
