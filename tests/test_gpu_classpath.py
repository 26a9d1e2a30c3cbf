"""The class path drop-in (NUFFT / SNM) on the GPU against the oracle."""
import numpy as np
import pytest

from helpers import workload
from oracle import matching

pytestmark = pytest.mark.gpu


def _settings(W, n_batch):
    from spinifel_amd.settings import SpinifelSettings
    s = SpinifelSettings()
    s.N_orientations = len(W.ref_quats)
    s.N_batch_size = n_batch
    s.reduced_det_shape = tuple(W.det_shape)
    s.nufft_eps = 1e-6
    return s


@pytest.mark.parametrize("precision,single", [("fp32", True), ("fp32", False), ("bf16", False)])
def test_snm_matches_oracle_class_path(precision, single):
    # single: use_single_prec (plain FP32 argmin); otherwise the FP32 form's
    # candidates are rescored in FP64, the reference's float64 default
    from spinifel_amd.classes import NUFFT, SNM
    W = workload("m41_1k_px")
    print(f"case: class path M=41, {len(W.ref_quats)} orientations, batches of 500, {precision}, "
          f"use_single_prec={single}", flush=True)
    s = _settings(W, 500)
    s.use_single_prec = single
    nufft = NUFFT(s, W.ppr, W.pdr, ref_orientations=W.ref_quats)
    snm = SNM(s, W.images, W.ppr, W.pdr, nufft, precision=precision)
    got = snm.slicing_and_match(W.ac)
    ora = matching.snm_slicing_and_match(W.ac, W.images, W.ppr, W.pdr, W.ref_quats, 500,
                                         return_details=True)
    np.testing.assert_array_equal(snm.last["index"], ora["index"])
    np.testing.assert_array_equal(got, W.ref_quats[ora["index"]])
    np.testing.assert_allclose(snm.last["mult"], ora["mult"], rtol=1e-5)
    # a second call reuses the resident images and the plan
    np.testing.assert_array_equal(snm.slicing_and_match(W.ac), got)


def test_nufft_forward_and_empty():
    import torch
    from spinifel_amd.classes import NUFFT, SNM
    W = workload("m41_1k_px")
    s = _settings(W, 500)
    nufft = NUFFT(s, W.ppr, W.pdr, ref_orientations=W.ref_quats)
    N = 3 * nufft.N_pixels
    out = nufft.forward(W.ac, 10, 13, 1, True, N).double().cpu().numpy()
    ref = matching.model_slices(W.ac, W.ref_quats[10:13], W.ppr, W.pdr.max(), 1, None,
                                1e-12).ravel()
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-5
    with pytest.raises(ValueError):
        nufft.forward(W.ac, 10, 13, 1, True, N + 1)
    snm = SNM(s, W.images[:0], W.ppr, W.pdr, nufft)
    assert snm.slicing_and_match(W.ac).shape == (0, 4)
