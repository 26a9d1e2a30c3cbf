"""GPU parity of the type-2 NUFFT slicing (libmtip) against the oracle:
FP64 direct NUDFT on subsamples (rel-L2 <= 1e-5, the stated bar) and the
oracle's ES-kernel NUFFT at eps 1e-12."""
import numpy as np
import pytest

from helpers import normalized_points, nudft_real, rel_l2, workload
from oracle import es_nufft, matching
from spinifel_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5   # slice relative L2 error vs the FP64 direct NUDFT


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    from spinifel_amd import engine
    engine.require_gpu()
    return torch.device("cuda", 0)


def _gpu_slices(ac, quats, ppr, E, eps=1e-6, os_=1, tile=(16, 16)):
    import torch
    from spinifel_amd import engine
    order = engine.tiling(ppr.shape[1:], tile) if tile else None
    plan = engine.NufftPlan(ac.shape[0], eps, 0)
    plan.set_model(ac)
    q = torch.as_tensor(np.ascontiguousarray(ppr.reshape(3, -1)))
    r = plan.slices(torch.as_tensor(quats), q, np.pi / (E * os_), order=order)
    assert int(r["bad"].item()) == 0
    n_p = q.shape[1]
    return r["s32"][:, :n_p].double().cpu().numpy(), r


@pytest.mark.parametrize("tile", [(16, 16), (8, 16), (8, 8), None])
def test_slices_m41_vs_nudft(torch_dev, tile):
    """16 x 16, 8 x 16 and 8 x 8 pixel tiles (LDS-staged boxes) and
    index-order runs of 256 pixels (boxes too big for LDS on this detector:
    direct-load fallback)."""
    print(f"case: M=41 1x32x32, 64 orientations, tile={tile}", flush=True)
    ac = synth.make_model(41, 300, 0)
    ppr, pdr = synth.detector(1, 32, 32)
    E = pdr.max()
    quats = synth.random_quaternions(64, np.random.default_rng(5))
    s, r = _gpu_slices(ac, quats, ppr, E, tile=tile)
    H, K, L = normalized_points(quats, ppr, E)
    rng = np.random.default_rng(0)
    sub = rng.choice(H.size, 3000, replace=False)
    ref = nudft_real(ac, H[sub], K[sub], L[sub])
    err = rel_l2(s.ravel()[sub], ref)
    print("rel_l2 vs NUDFT", err)
    assert err <= TOL
    # whole set against the oracle's FINUFFT-method transform at 1e-12
    full = matching.model_slices(ac, quats, ppr, E, 1, None, 1e-12)
    assert rel_l2(s, full) <= TOL
    # FP64 row statistics of the written values
    np.testing.assert_allclose(r["row_sum"].cpu().numpy(), s.sum(1), rtol=1e-10)
    np.testing.assert_allclose(r["row_sumsq"].cpu().numpy(), (s * s).sum(1), rtol=1e-10)


def test_slices_m81_16k_px_vs_nudft(torch_dev):
    print("case: M=81 4x32x128, 6 orientations", flush=True)
    ac = synth.make_model(81, 800, 0)
    ppr, pdr = synth.detector(4, 32, 128)
    E = pdr.max()
    quats = synth.random_quaternions(6, np.random.default_rng(7))
    s, _ = _gpu_slices(ac, quats, ppr, E)
    H, K, L = normalized_points(quats, ppr, E)
    sub = np.random.default_rng(1).choice(H.size, 1500, replace=False)
    ref = nudft_real(ac, H[sub], K[sub], L[sub])
    err = rel_l2(s.ravel()[sub], ref)
    print("rel_l2 vs NUDFT (M=81)", err)
    assert err <= TOL


def test_slices_m121_16k_px_small_tiles(torch_dev):
    """M = 121 (n = 250 fine grid): 16 x 16 boxes exceed LDS, 8 x 8 tiles stage."""
    print("case: M=121 4x32x128, 4 orientations, 8x8 and 16x16 tiles", flush=True)
    ac = synth.make_model(121, 1200, 1)
    ppr, pdr = synth.detector(4, 32, 128)
    E = pdr.max()
    quats = synth.random_quaternions(4, np.random.default_rng(9))
    s8, _ = _gpu_slices(ac, quats, ppr, E, tile=(8, 8))
    s16, _ = _gpu_slices(ac, quats, ppr, E, tile=(16, 16))
    H, K, L = normalized_points(quats, ppr, E)
    sub = np.random.default_rng(2).choice(H.size, 800, replace=False)
    ref = nudft_real(ac, H[sub], K[sub], L[sub])
    assert rel_l2(s8.ravel()[sub], ref) <= TOL
    assert rel_l2(s16.ravel()[sub], ref) <= TOL
    # identical arithmetic per point whatever the tiling
    np.testing.assert_array_equal(s8, s16)


@pytest.mark.parametrize("M", [20, 33])
def test_even_odd_and_nonsymmetric_model(torch_dev, M):
    """Re of the transform of a non-centrosymmetric model = transform of its
    symmetric part; even M has the unpaired -M/2 mode."""
    print(f"case: random model M={M}", flush=True)
    rng = np.random.default_rng(M)
    ac = rng.standard_normal((M, M, M)).astype(np.float32)
    x = rng.uniform(-np.pi, np.pi, (3, 2000))
    from spinifel_amd import engine
    plan = engine.NufftPlan(M, 1e-6, 0)
    plan.set_model(ac)
    out = plan.execute(x[0], x[1], x[2]).double().cpu().numpy()
    ref = nudft_real(ac, x[0], x[1], x[2])
    err = rel_l2(out, ref)
    print("rel_l2", err)
    assert err <= TOL


def test_execute_wraps_periodic_domain(torch_dev):
    """Points in [pi, 3pi) equal their 2pi-shifted images (the transform is
    2pi-periodic in each coordinate)."""
    from spinifel_amd import engine
    ac = synth.make_model(21, 40, 3)
    plan = engine.NufftPlan(21, 1e-6, 0)
    plan.set_model(ac)
    rng = np.random.default_rng(2)
    x = rng.uniform(-np.pi, np.pi, (3, 500))
    a = plan.execute(x[0], x[1], x[2]).double().cpu().numpy()
    b = plan.execute(x[0] + 2 * np.pi, x[1] - 2 * np.pi, x[2]).double().cpu().numpy()
    assert rel_l2(b, a) <= 1e-5


def test_domain_assert(torch_dev):
    from spinifel_amd import engine
    ac = synth.make_model(21, 40, 3)
    plan = engine.NufftPlan(21, 1e-6, 0)
    plan.set_model(ac)
    with pytest.raises(AssertionError):
        plan.execute([0.0, 3.1 * np.pi], [0.0, 0.0], [0.0, 0.0])
    with pytest.raises(AssertionError):
        plan.execute([np.nan], [0.0], [0.0])
    # in-domain points are still fine afterwards
    v = plan.execute([0.1], [0.2], [0.3]).double().cpu().numpy()
    ref = nudft_real(ac, [0.1], [0.2], [0.3])
    assert abs(v[0] - ref[0]) <= 1e-5 * abs(ref[0])


def test_tolerances_reach_fp32_floor(torch_dev):
    """Wider kernels (tighter eps) converge to the FP32 floor."""
    from spinifel_amd import engine
    ac = synth.make_model(41, 300, 0)
    rng = np.random.default_rng(9)
    x = rng.uniform(-0.9 * np.pi, 0.9 * np.pi, (3, 2000))
    ref = nudft_real(ac, x[0], x[1], x[2])
    errs = {}
    for eps in (1e-3, 1e-6, 1e-9):
        plan = engine.NufftPlan(41, eps, 0)
        plan.set_model(ac)
        errs[eps] = rel_l2(plan.execute(x[0], x[1], x[2]).double().cpu().numpy(), ref)
    print(errs)
    assert errs[1e-3] <= 1e-2
    assert errs[1e-6] <= 1e-5
    assert errs[1e-9] <= 2e-6
