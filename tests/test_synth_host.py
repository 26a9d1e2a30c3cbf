"""Host-side pieces: synthetic workload generators, tile order, settings."""
import numpy as np
import pytest

from spinifel_amd import engine, settings as st, synth


def test_model_is_real_centrosymmetric_autocorrelation():
    ac = synth.make_model(21, 30, 0)
    assert ac.dtype == np.float32 and ac.shape == (21, 21, 21)
    np.testing.assert_array_equal(ac, ac[::-1, ::-1, ::-1])
    rho = synth.density(21, 30, synth.rng_for(0, synth.STREAM_ATOMS))
    # the centre of an autocorrelation is sum(rho^2), its maximum
    assert ac[10, 10, 10] == pytest.approx((rho ** 2).sum(), rel=1e-5)
    assert ac.max() == ac[10, 10, 10]
    np.testing.assert_array_equal(ac, synth.make_model(21, 30, 0))


def test_detector_geometry():
    ppr, pdr = synth.detector(4, 32, 128)
    assert ppr.shape == (3, 4, 32, 128) and pdr.shape == (4, 32, 128)
    q = np.linalg.norm(ppr, axis=0)
    np.testing.assert_allclose(pdr, q / 0.9)
    # Ewald sphere: |q + z_hat / lambda| = 1 / lambda
    np.testing.assert_allclose(np.linalg.norm(ppr + np.array([0, 0, 1.0])[:, None, None, None], axis=0),
                               1.0, rtol=1e-12)


def test_quaternions_and_images():
    q = synth.random_quaternions(1000, np.random.default_rng(0))
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0)
    assert (q[:, 0] >= 0).all()
    s = np.abs(np.random.default_rng(1).standard_normal((400, 64)))
    im = synth.poisson_images(s, 2.0, np.random.default_rng(2))
    assert im.dtype == np.float32
    assert abs(im.mean() - 2.0) < 0.1
    assert (im == np.round(im)).all()


@pytest.mark.parametrize("shape", [(4, 32, 128), (1, 32, 32), (1, 20, 13), (3, 7, 40)])
def test_tile_order_is_a_permutation(shape):
    o = engine.tile_order(shape)
    assert o.size % 256 == 0
    valid = o[o >= 0]
    np.testing.assert_array_equal(np.sort(valid), np.arange(int(np.prod(shape))))
    # each run of 256 stays inside one 16 x 16 block of one panel
    rows, cols = shape[-2:]
    for run in o.reshape(-1, 256):
        v = run[run >= 0]
        if v.size:
            p, r, c = np.unravel_index(v, (int(np.prod(shape[:-2])), rows, cols))
            assert len(set(p)) == 1 and r.max() - r.min() < 16 and c.max() - c.min() < 16
    assert engine.tile_order((100,)) is None


def test_tile_order_groups_neighbours():
    # inside a full 16 x 16 tile, every run of 4 slots is a 2 x 2 pixel block
    shape = (4, 32, 128)
    o = engine.tile_order(shape)
    for run in o.reshape(-1, 4):
        p, r, c = np.unravel_index(run, shape)
        assert len(set(p)) == 1
        assert r.min() % 2 == 0 and c.min() % 2 == 0
        np.testing.assert_array_equal(r - r.min(), [0, 0, 1, 1])
        np.testing.assert_array_equal(c - c.min(), [0, 1, 0, 1])
    # group=None keeps row-major tiles
    o_row = engine.tile_order(shape, group=None)
    np.testing.assert_array_equal(o_row[:16], np.arange(16))
    np.testing.assert_array_equal(np.sort(o_row), np.sort(o))


def test_spatial_order_groups_cubes():
    import torch
    rng = np.random.default_rng(5)
    pts = torch.as_tensor(rng.uniform(-3, 3, (3, 5000)))
    cell = 0.5
    perm = engine.spatial_order(pts[0], pts[1], pts[2], cell)
    assert sorted(perm.tolist()) == list(range(5000))
    # every cube's points form one contiguous run of the order
    cube = ((pts + 3 * np.pi) / cell).floor().long()
    key = (cube[0] * 1024 + cube[1]) * 1024 + cube[2]
    k = key[perm].numpy()
    starts = np.flatnonzero(np.r_[True, k[1:] != k[:-1]])
    assert len(starts) == len(np.unique(k))
    # Morton order: the first 8 cubes of a 2x2x2 block come before the next block
    c = cube[:, perm].numpy()
    blk = (c[0] // 2) * 1_000_000 + (c[1] // 2) * 1000 + (c[2] // 2)
    b_starts = np.flatnonzero(np.r_[True, blk[1:] != blk[:-1]])
    assert len(b_starts) == len(np.unique(blk))


def test_settings_shim():
    s = st.SpinifelSettings()
    assert s.Mquat == 20 and s.M == 81
    assert st.SpinifelSettings(oversampling=2).M == 161
