"""GPU parity of the type-1 NUFFT (the adjoint of slicing, libmtip) against
the oracle: the FP64 brute-force sum and the oracle's ES type 1 at eps 1e-12
(itself pinned to the brute force), rel-L2 <= 1e-5 -- the same bar as the
slices.  Also the adjoint identity between the GPU's own type-2 and type-1
kernels, accumulation over several spread calls, support / scale handling
and the domain check."""
import ctypes

import numpy as np
import pytest

from helpers import normalized_points, rel_l2
from oracle import es_nufft
from spinifel_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    from spinifel_amd import engine
    engine.require_gpu()
    return torch.device("cuda", 0)


def _geometry(M_det=(1, 32, 32), n_o=64, seed=5):
    ppr, pdr = synth.detector(*M_det)
    E = float(pdr.max())
    quats = synth.random_quaternions(n_o, np.random.default_rng(seed))
    return ppr, E, quats


def _oracle_adjoint(vals, H, K, L, M, support=None):
    """spinifel's adjoint (real part, support, / M^3) with the ES type 1 at 1e-12."""
    return es_nufft.adjoint(np.asarray(vals, np.float64).ravel(), H, K, L, M, True, support)


@pytest.mark.parametrize("spatial", [False, True])
@pytest.mark.parametrize("M", [16, 17])
def test_points_vs_brute_force(torch_dev, M, spatial):
    """Scattered points (in caller order the boxes are too large for LDS: the
    direct FP64-atomic path; in spatial order, compact LDS boxes) and points
    beyond pi (periodic wrap)."""
    from spinifel_amd import engine
    print(f"case: adjoint_points M={M} 4000 random points spatial={spatial}", flush=True)
    rng = np.random.default_rng(M)
    x = rng.uniform(-2.9 * np.pi, 2.9 * np.pi, (3, 4000))
    v = rng.standard_normal(4000).astype(np.float32)
    plan = engine.NufftPlan(M, 1e-6, 0)
    out = plan.adjoint_points(*x, values=v, spatial=spatial).cpu().numpy()
    ref = es_nufft.nudft3d1(*x, v.astype(np.float64), M, 1).real / M ** 3
    assert out.shape == (M,) * 3
    assert rel_l2(out, ref) < TOL


@pytest.mark.parametrize("M,tile", [(41, (16, 16)), (41, (8, 16)), (41, None), (82, (16, 16)),
                                    (82, (8, 8))])
def test_slices_vs_oracle(torch_dev, M, tile):
    """Detector tiles of one orientation per workgroup (LDS box path); M = 82
    is the upsampled edge of the convolution kernel for M = 41 (values = ones,
    as Merge's nuvect)."""
    import torch
    from spinifel_amd import engine
    print(f"case: adjoint_slices M={M} tile={tile}", flush=True)
    ppr, E, quats = _geometry()
    N_p = int(np.prod(ppr.shape[1:]))
    plan = engine.NufftPlan(M, 1e-6, 0)
    order = engine.tiling(ppr.shape[1:], tile) if tile else None
    q = torch.as_tensor(np.ascontiguousarray(ppr.reshape(3, -1)))
    ones = M % 2 == 0
    vals = None if ones else \
        np.random.default_rng(1).standard_normal((len(quats), N_p)).astype(np.float32)
    out = plan.adjoint_slices(torch.as_tensor(quats), q, np.pi / E,
                              values=None if ones else torch.as_tensor(vals), order=order)
    H, K, L = normalized_points(quats, ppr, E)
    ref = _oracle_adjoint(np.ones(H.size) if ones else vals, H, K, L, M)
    assert rel_l2(out.cpu().numpy(), ref) < TOL


def test_m81_16k_px_vs_oracle(torch_dev):
    """The matching workload's grid and detector (M = 81, 128 x 128 pixels),
    32 orientations, values as images."""
    import torch
    from spinifel_amd import engine
    print("case: adjoint_slices M=81 1x128x128, 32 orientations", flush=True)
    ppr, E, quats = _geometry((1, 128, 128), 32, 9)
    plan = engine.NufftPlan(81, 1e-6, 0)
    q = torch.as_tensor(np.ascontiguousarray(ppr.reshape(3, -1)))
    vals = np.random.default_rng(2).uniform(0, 4, (32, q.shape[1])).astype(np.float32)
    out = plan.adjoint_slices(torch.as_tensor(quats), q, np.pi / E, values=torch.as_tensor(vals),
                              order=engine.tiling(ppr.shape[1:], (16, 16)))
    H, K, L = normalized_points(quats, ppr, E)
    assert rel_l2(out.cpu().numpy(), _oracle_adjoint(vals, H, K, L, 81)) < TOL


def test_adjoint_identity_with_slices(torch_dev):
    """<slices(f), c> = <f, M^3 adjoint(c)> between the GPU's own two kernels
    (real f and c, so real parts suffice)."""
    import torch
    from spinifel_amd import engine
    print("case: adjoint identity M=41", flush=True)
    M = 41
    ppr, E, quats = _geometry(n_o=32)
    f = synth.make_model(M, 300, 3)
    plan = engine.NufftPlan(M, 1e-6, 0)
    plan.set_model(f)
    q = torch.as_tensor(np.ascontiguousarray(ppr.reshape(3, -1)))
    order = engine.tiling(ppr.shape[1:], (16, 16))
    N_p = q.shape[1]
    s = plan.slices(torch.as_tensor(quats), q, np.pi / E, order=order)["s32"][:, :N_p]
    c = torch.as_tensor(np.random.default_rng(4).standard_normal((32, N_p)).astype(np.float32))
    a = plan.adjoint_slices(torch.as_tensor(quats), q, np.pi / E, values=c, order=order,
                            out_scale=1.0)
    fs = 0.5 * (f + f[::-1, ::-1, ::-1])       # slices see the symmetrised model
    lhs = float((s.double().cpu() * c.double()).sum())
    rhs = float((torch.as_tensor(fs.copy()) * a.cpu()).sum())
    bound = float(np.linalg.norm(s.double().cpu().numpy()) * np.linalg.norm(c.numpy()))
    assert abs(lhs - rhs) <= 1e-5 * bound


def test_spread_calls_accumulate(torch_dev):
    """begin + two spread calls (halves of the orientations) + finish equals
    one call over all, through the C ABI; support and scale applied at finish."""
    import torch
    from spinifel_amd import _lib, engine
    print("case: accumulate over two spread calls M=33", flush=True)
    M = 33
    ppr, E, quats = _geometry(n_o=40)
    plan = engine.NufftPlan(M, 1e-6, 0)
    lib = _lib.load()
    dev = torch.device("cuda", 0)
    qt = torch.as_tensor(quats).to(dev)
    q = torch.as_tensor(np.ascontiguousarray(ppr.reshape(3, -1))).to(dev)
    N_p = q.shape[1]
    vals = torch.as_tensor(np.random.default_rng(6).standard_normal((40, N_p)).astype(
        np.float32)).to(dev)
    sup = torch.as_tensor((np.random.default_rng(7).uniform(size=(M,) * 3) > 0.5).astype(
        np.float64)).to(dev)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    out = torch.empty((M,) * 3, dtype=torch.float64, device=dev)
    st = engine._stream()
    p = engine._ptr
    _lib.check(lib.mtip_plan_spread_begin(plan._h, st))
    for lo, hi in ((0, 25), (25, 40)):
        _lib.check(lib.mtip_plan_spread_slices(
            plan._h, ctypes.c_void_p(qt[lo:].data_ptr()), hi - lo, p(q), N_p, None, 0, 0,
            float(np.pi / E), ctypes.c_void_p(vals[lo:].data_ptr()), N_p, p(bad), st))
    _lib.check(lib.mtip_plan_spread_finish(plan._h, p(sup), 2.0, p(out), st))
    one = plan.adjoint_slices(qt, q, np.pi / E, values=vals, support=sup, out_scale=2.0)
    torch.cuda.synchronize()
    assert int(bad.item()) == 0
    assert rel_l2(out.cpu().numpy(), one.cpu().numpy()) < 1e-12
    assert float(out[sup == 0].abs().max()) == 0.0


def test_finish_before_begin_and_domain(torch_dev):
    from spinifel_amd import _lib, engine
    plan = engine.NufftPlan(12, 1e-6, 0)
    out = engine.torch.empty((12,) * 3, dtype=engine.torch.float64, device="cuda")
    rc = _lib.load().mtip_plan_spread_finish(plan._h, None, 1.0, engine._ptr(out), engine._stream())
    assert rc == _lib.MTIP_ERR_STATE
    x = np.array([[0.1, 3.5 * np.pi], [0.0, 0.0], [0.0, 0.0]])
    with pytest.raises(AssertionError):
        plan.adjoint_points(*x)


def test_dropin_adjoint_matches_oracle(torch_dev):
    """spinifel_amd.autocorrelation.adjoint keeps the reference signature
    (spinifel/autocorrelation.py:97) and result: float64 (M, M, M)."""
    from spinifel_amd import autocorrelation as ac_mod
    print("case: drop-in adjoint M=21", flush=True)
    M = 21
    ppr, E, quats = _geometry(n_o=8)
    H, K, L = normalized_points(quats, ppr, E)
    v = np.random.default_rng(8).standard_normal(H.size)
    sup = (np.random.default_rng(9).uniform(size=(M,) * 3) > 0.2).astype(np.float64)
    out = ac_mod.adjoint(v.astype(np.complex128), H, K, L, M, use_recip_sym=True, support=sup)
    ref = _oracle_adjoint(v.astype(np.float32), H, K, L, M, sup)
    assert isinstance(out, np.ndarray) and out.dtype == np.float64 and out.shape == (M,) * 3
    assert rel_l2(out, ref) < TOL


def test_points_spatial_order_large(torch_dev):
    """2^20 points (the size from which the spatial order is automatic): the
    same sum as the caller's order, up to the FP32 accumulation in LDS."""
    import torch
    from spinifel_amd import engine
    print("case: adjoint_points M=33 2^20 points, spatial order vs caller order", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    n = 1 << 20
    x = (torch.rand((3, n), generator=g, device="cuda", dtype=torch.float64) * 2 - 1) * 2.5
    v = torch.randn(n, generator=g, device="cuda", dtype=torch.float32)
    plan = engine.NufftPlan(33, 1e-6, 0)
    a = plan.adjoint_points(*x, values=v, spatial=False)
    b = plan.adjoint_points(*x, values=v)          # 'auto' sorts at this size
    rel = float((a - b).norm() / a.norm())
    print("rel-L2 spatial vs caller order", rel, flush=True)
    assert rel < 1e-6
