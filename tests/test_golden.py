"""Committed golden case (tests/golden/m21_16x16.npz, written by
tests/golden/make_golden.py from the oracle): the oracle reproduces it (CPU)
and the GPU drop-in path matches it."""
import os

import numpy as np
import pytest

from helpers import rel_l2
from oracle import matching

GOLD = os.path.join(os.path.dirname(__file__), "golden", "m21_16x16.npz")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD))


def test_oracle_reproduces_golden(gold):
    r = matching.slicing_and_match(gold["ac"], gold["images"], gold["ppr"], gold["pdr"],
                                   gold["ref_quats"], return_details=True)
    np.testing.assert_array_equal(r["index"], gold["index"])
    assert r["alpha"] == pytest.approx(float(gold["alpha"]), rel=1e-12)
    np.testing.assert_allclose(r["min_dist"], gold["min_dist"], rtol=1e-10)
    ms = r["model_slices"] / r["alpha"]
    np.testing.assert_allclose(ms.ravel()[gold["slice_flat_idx"]], gold["slice_values"], rtol=1e-10,
                               atol=1e-12 * np.abs(gold["slice_values"]).max())


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp8"])
def test_gpu_matches_golden(gold, precision):
    from spinifel_amd import orientation_matching, settings as st, slicing
    N_o = gold["ref_quats"].shape[0]
    det = gold["images"].shape[1:]
    s = st.SpinifelSettings(N_orientations=N_o, N_batch_size=N_o, reduced_det_shape=det)
    r = orientation_matching.slicing_and_match(gold["ac"], gold["images"], gold["ppr"],
                                               gold["pdr"], gold["ref_quats"], settings=s,
                                               precision=precision, return_details=True)
    np.testing.assert_array_equal(r["index"], gold["index"])
    np.testing.assert_allclose(r["alpha"][0], gold["alpha"], rtol=1e-5)
    ms = slicing.gen_model_slices(gold["ac"], gold["ref_quats"], gold["ppr"],
                                  float(gold["pdr"].max()), 1, None, int(np.prod(det)))
    assert rel_l2(ms[gold["slice_flat_idx"]], gold["slice_values"]) <= 1e-5
