"""Committed golden case at the quickstart scale (tests/golden/m81_16k_px.npz,
written by tests/golden/make_m81_golden.py from the oracle): M=81 model,
4 x 32 x 128 = 16,384 px, 10,000 reference orientations, 1,000 Poisson
images, oracle NUFFT eps 1e-12 and FP64 distances (the reference CPU path,
spinifel/sequential/orientation_matching.py:189-256 with
spinifel/sequential/nearest_neighbor.py:131-135).

CPU: the oracle reproduces the fixture's slice subsample and the FP64
distances of a few images' candidates.  GPU: the drop-in ``match`` called
through the settings singleton (spinifel/mpi/orientation_matching.py:8-16)
returns the oracle's orientation for every image in the plain FP32 form and
in the BF16 / FP8 forms with FP32 + FP64 rescoring; the std ratio agrees to
1e-5; ``gen_model_slices_batch`` agrees with the oracle's slices to a
relative L2 of 1e-5 (the NUFFT bar) on the 4,000-value subsample.

The smallest relative margin between an image's best and second-best FP64
squared distance in this fixture is 2.1e-5 (image 103; image 938 has
2.9e-5, every other image more than 5e-5), against an FP32 distance error
of about 5e-6 relative, so no image is expected to flip in FP32; a flip in
any form fails the test.
"""
import os

import numpy as np
import pytest

from helpers import rel_l2
from oracle import es_nufft, geometry

GOLD = os.path.join(os.path.dirname(__file__), "golden", "m81_16k_px.npz")


def load_gold():
    g = dict(np.load(GOLD))
    M = int(g["M"])
    half = g.pop("ac_half")
    # centrosymmetric model: a point reflection reverses the flattened index
    g["ac"] = np.concatenate([half, half[:-1][::-1]]).reshape(M, M, M)
    g["images"] = g["images"].astype(np.float32)
    return g


@pytest.fixture(scope="module")
def gold():
    return load_gold()


def test_fixture_is_consistent(gold):
    ac = gold["ac"]
    assert np.array_equal(ac, ac[::-1, ::-1, ::-1])
    assert gold["ref_quats"].shape == (10000, 4) and gold["images"].shape == (1000, 4, 32, 128)
    np.testing.assert_array_equal(gold["top_idx"][:, 0], gold["index"])
    np.testing.assert_allclose(np.sqrt(gold["top_d2"][:, 0]), gold["min_dist"], rtol=1e-9)
    assert np.all(np.diff(gold["top_d2"], axis=1) >= 0)


def test_oracle_reproduces_slice_subsample(gold):
    N_p = 16384
    flat = gold["slice_flat_idx"]
    o, p = flat // N_p, flat % N_p
    E = float(gold["pdr"].max())
    rot = geometry.ref_rotmat(gold["ref_quats"][o])
    q = gold["ppr"].reshape(3, -1)[:, p]
    hkl = np.einsum("nij,jn->in", rot, q) * np.pi / E
    got = es_nufft.nufft3d2(hkl[0], hkl[1], hkl[2], gold["ac"].astype(np.float64), -1, 1e-12).real
    assert rel_l2(got, gold["slice_values"]) < 1e-10


def test_oracle_reproduces_candidate_distances(gold):
    # FP64 squared distances ||alpha s_n - d_i||^2 of 3 images' 8 candidates
    imgs = [0, 103, 938]
    E = float(gold["pdr"].max())
    cand = gold["top_idx"][imgs]
    quats = gold["ref_quats"][cand.ravel()]
    H, K, L = geometry.gen_nonuniform_normalized_positions(quats, gold["ppr"], E, 1)
    s = es_nufft.nufft3d2(H, K, L, gold["ac"].astype(np.float64), -1, 1e-12).real
    s = s.reshape(len(imgs), 8, -1) * float(gold["alpha"])
    d = gold["images"][imgs].reshape(len(imgs), 1, -1).astype(np.float64)
    d2 = ((s - d) ** 2).sum(-1)
    np.testing.assert_allclose(d2, gold["top_d2"][imgs], rtol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp32_plain", "fp32", "bf16", "fp8"])
def test_dropin_match_through_settings_equals_oracle(gold, form):
    # fp32_plain: use_single_prec, the plain FP32 argmin (K = 1, no rescoring);
    # fp32: the default float64 semantics, FP32 candidates rescored in FP64
    from spinifel_amd import orientation_matching, settings as st

    precision = form.split("_")[0]
    s = st.settings
    saved = dict(vars(s))
    try:
        s.N_orientations = 10000
        s.N_batch_size = int(gold["N_batch_size"])
        s.reduced_det_shape = (4, 32, 128)
        s.oversampling = 1
        s.use_single_prec = form == "fp32_plain"
        s.precision = precision
        s.topk = 0
        s.rescore = None
        print(f"m81 golden: match() {form}", flush=True)
        got = orientation_matching.match(gold["ac"], gold["images"], gold["ppr"], gold["pdr"],
                                         gold["ref_quats"])
        det = orientation_matching.slicing_and_match(gold["ac"], gold["images"], gold["ppr"],
                                                     gold["pdr"], gold["ref_quats"],
                                                     return_details=True)
    finally:
        for k, v in saved.items():
            setattr(s, k, v)
    ref = gold["ref_quats"][gold["index"]]
    same = np.all(got == ref, axis=1)
    d = gold["top_d2"]
    margin = (d[:, 1] - d[:, 0]) / d[:, 0]
    bad = np.flatnonzero(~same)
    print(f"m81 golden {form}: {same.sum()} / {same.size} identical; "
          f"flipped images {bad.tolist()} with FP64 margins {margin[bad].tolist()}", flush=True)
    assert same.all()
    np.testing.assert_array_equal(det["index"], gold["index"])
    np.testing.assert_allclose(det["alpha"][0], gold["alpha"], rtol=1e-5)


@pytest.mark.gpu
def test_gen_model_slices_batch_equals_oracle(gold):
    from spinifel_amd import slicing

    E = float(gold["pdr"].max())
    out = slicing.gen_model_slices_batch(gold["ac"], gold["ref_quats"], gold["ppr"], E, 1, None,
                                         16384, batch_size=1000)
    assert out.dtype == np.float64 and out.shape == (10000 * 16384,)
    err = rel_l2(out[gold["slice_flat_idx"]], gold["slice_values"])
    print(f"m81 golden: gen_model_slices_batch rel-L2 vs oracle eps 1e-12 = {err:.2e}", flush=True)
    assert err <= 1e-5
    with pytest.raises(AssertionError):
        slicing.gen_model_slices_batch(gold["ac"], gold["ref_quats"][:999], gold["ppr"], E, 1,
                                       None, 16384, batch_size=100)
