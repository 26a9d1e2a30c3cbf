"""GPU parity of the merge step (Merge.solve_ac) against the oracle's FP64
restatement of spinifel/sequential/autocorrelation.py:35-225: the operator
pieces one by one, the CG iteration, and whole solves with and without an
estimate (same iteration count, rel-L2 <= 1e-4 on ac -- the transforms run
at eps 1e-6 against the oracle's 1e-12).  At M = 81 the operator is checked
for symmetry and the solve for residual reduction."""
from types import SimpleNamespace

import numpy as np
import pytest

from helpers import normalized_points, rel_l2
from oracle import es_nufft
from oracle import merge as omerge
from spinifel_amd import synth

pytestmark = pytest.mark.gpu

TOL_AC = 1e-4


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    from spinifel_amd import engine
    engine.require_gpu()
    return torch.device("cuda", 0)


def _settings(M, maxiter, os_=1):
    return SimpleNamespace(M=M, M_ups=2 * M, oversampling=os_, solve_ac_maxiter=maxiter,
                           nufft_eps=1e-6, device=0)


@pytest.fixture(scope="module")
def problem():
    """17^3 model, 24 images of 16 x 16 pixels at random orientations,
    images = Re of the model's transform (oracle direct sum) + noise."""
    M = 17
    ppr, pdr = synth.detector(1, 16, 16)
    E = float(pdr.max())
    rng = np.random.default_rng(0)
    quats = synth.random_quaternions(24, rng)
    model = synth.make_model(M, 60, 1)
    H, K, L = normalized_points(quats, ppr, E)
    img = es_nufft.nudft3d2(H, K, L, model, -1).real
    img = img + 0.01 * np.abs(img).mean() * rng.standard_normal(img.shape)
    return dict(M=M, ppr=ppr, pdr=pdr, E=E, quats=quats, model=model, H=H, K=K, L=L,
                images=img.reshape(24, 1, 16, 16))


def test_operators_vs_oracle(torch_dev, problem):
    import torch
    from spinifel_amd.merge import Merge
    print("case: merge operators M=17", flush=True)
    p = problem
    M = p["M"]
    mg = Merge(_settings(M, 5), p["images"], p["ppr"], p["pdr"])
    rng = np.random.default_rng(3)
    u = rng.standard_normal(M ** 3)
    s = (rng.uniform(size=(M,) * 3) > 0.2).astype(np.float64)
    ut = torch.as_tensor(u, device=torch_dev)
    st = torch.as_tensor(s, device=torch_dev)
    # Fourier regulariser: FFTs only, FP64 both sides
    A = omerge.antisupport(M, 1)
    np.testing.assert_allclose(mg.fourier_reg(ut, st).cpu().numpy(),
                               omerge.fourier_reg(u, s, A, M), rtol=0, atol=1e-12)
    # convolution with the kernel of the GPU's type 1 vs the oracle's
    W, d = mg.setup_linops(p["quats"], st, torch.zeros(M ** 3, dtype=torch.float64,
                                                        device=torch_dev))
    _, F = omerge.conv_kernel(p["H"], p["K"], p["L"], M, 2 * M)
    x0 = np.zeros(M ** 3)
    d_ref = omerge.data_term(p["images"].reshape(-1), p["H"], p["K"], p["L"], s, x0, M,
                             mg.rlambda)
    assert rel_l2(d.cpu().numpy(), d_ref) < 1e-5
    ref_W = (omerge.core_problem_convolution(u, M, F, 2 * M, s) + mg.rlambda * u
             + 1e3 * omerge.fourier_reg(u, s, A, M))
    assert rel_l2(W(ut).cpu().numpy(), ref_W) < 1e-5


def test_gaussian_filter_and_cg_vs_oracle(torch_dev):
    import torch
    from scipy.ndimage import gaussian_filter
    from spinifel_amd import merge
    a = np.random.default_rng(4).standard_normal((11, 12, 13))
    np.testing.assert_allclose(merge.gaussian_filter(torch.as_tensor(a, device=torch_dev)).cpu(),
                               gaussian_filter(a, 0.5), rtol=1e-12, atol=1e-14)
    rng = np.random.default_rng(5)
    Q = rng.standard_normal((50, 50))
    Am = Q @ Q.T + 50 * np.eye(50)
    b, x0 = rng.standard_normal(50), rng.standard_normal(50)
    At = torch.as_tensor(Am, device=torch_dev)
    for maxiter in (4, 200):
        xr, info_r, it_r = omerge.cg(lambda v: Am @ v, b, x0, maxiter)
        for solver in (merge.cg, merge.cg_fused):     # torch ops, and libmtip's kernels
            x, info, it = solver(lambda v: At @ v, torch.as_tensor(b, device=torch_dev),
                                 torch.as_tensor(x0, device=torch_dev), maxiter)
            assert (info, it) == (info_r, it_r)
            np.testing.assert_allclose(x.cpu().numpy(), xr, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("with_estimate", [False, True])
def test_solve_ac_vs_oracle(torch_dev, problem, with_estimate):
    from spinifel_amd.merge import Merge
    print(f"case: solve_ac M=17 estimate={with_estimate}", flush=True)
    p = problem
    M = p["M"]
    maxiter = 6
    est = None
    if with_estimate:
        est = np.zeros((M,) * 3)
        c = M // 2
        est[c - 4:c + 5, c - 4:c + 5, c - 4:c + 5] = np.abs(p["model"][c - 4:c + 5, c - 4:c + 5,
                                                                        c - 4:c + 5])
    mg = Merge(_settings(M, maxiter), p["images"], p["ppr"], p["pdr"])
    ac = mg.solve_ac(0, p["quats"], None if est is None else est.copy())
    ref, det = omerge.solve_ac(p["images"], p["H"], p["K"], p["L"], M, 2 * M, 1, maxiter,
                               ac_estimate=est, return_details=True)
    assert ac.shape == (M,) * 3 and ac.dtype == np.float64
    assert mg.last["iterations"] == det["iterations"] and mg.callback.counter == det["iterations"]
    np.testing.assert_array_equal(mg.last["support"].cpu().numpy(), det["support"])
    assert rel_l2(ac, ref) < TOL_AC


def test_m81_operator_symmetric_and_solve_converges(torch_dev):
    """The default grid (M = 81, M_ups = 162) with 200 images of 64 x 64
    pixels: W is symmetric (<u, W v> = <W u, v>) and positive, and 100 CG
    iterations (solve_ac_maxiter) bring the residual well below ||d||."""
    import torch
    from spinifel_amd.merge import Merge
    from spinifel_amd.settings import SpinifelSettings
    print("case: merge M=81 200 images 64x64", flush=True)
    ppr, pdr = synth.detector(1, 64, 64)
    rng = np.random.default_rng(6)
    quats = synth.random_quaternions(200, rng)
    images = rng.uniform(0, 2, (200, 1, 64, 64))
    settings = SpinifelSettings(reduced_det_shape=(1, 64, 64))
    mg = Merge(settings, images, ppr, pdr)
    M = mg.M
    assert (M, mg.M_ups) == (81, 162)
    ones = torch.ones((M,) * 3, dtype=torch.float64, device=torch_dev)
    W, d = mg.setup_linops(quats, ones, torch.zeros(M ** 3, dtype=torch.float64, device=torch_dev))
    u = torch.as_tensor(rng.standard_normal(M ** 3), device=torch_dev)
    v = torch.as_tensor(rng.standard_normal(M ** 3), device=torch_dev)
    uWv, Wuv = float(u @ W(v)), float(W(u) @ v)
    assert abs(uWv - Wuv) <= 1e-6 * float(torch.linalg.norm(W(u)) * torch.linalg.norm(v))
    assert float(u @ W(u)) > 0
    ac = mg.solve_ac(0, quats)
    assert np.all(np.isfinite(ac))
    it, info = mg.last["iterations"], mg.last["info"]
    r = mg.last["W"](mg.last["ac"].reshape(-1)) - mg.last["d"]
    rel = float(torch.linalg.norm(r)) / float(torch.linalg.norm(mg.last["d"]))
    print(f"iterations {it} info {info} relative residual {rel:.3e}", flush=True)
    assert it <= 100 and (info == 0) == (rel <= 1.01e-5)
    assert rel < 5e-2        # 1.0e-2 measured on this (inconsistent, random) data
