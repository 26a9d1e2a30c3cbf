"""Batching form on the GPU: grouped top-K against a brute force, the
hierarchy against the flat search over its finest level, HIP-graph replays
against eager passes, and per-batch scaling."""
import numpy as np
import pytest
import torch

from helpers import oracle_slicer
from spinifel_amd import _lib, engine, hierarchy, synth
from spinifel_amd.engine import _ptr

pytestmark = pytest.mark.gpu


def _bf16_round(x):
    """float32 -> nearest-even bf16 values as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_group_topk_equals_brute_force():
    rng = np.random.default_rng(0)
    dev = torch.device("cuda", 0)
    n_c, n_f, N_i, P, K = 23, 700, 333, 320, 4
    S32 = rng.standard_normal((n_f, P)).astype(np.float32)
    D32 = rng.standard_normal((N_i, P)).astype(np.float32)
    ld = engine.padded(P)
    Sd = torch.zeros((n_f, ld), dtype=torch.float32, device=dev)
    Sd[:, :P] = torch.as_tensor(S32)
    Dd = torch.zeros((N_i, ld), dtype=torch.float32, device=dev)
    Dd[:, :P] = torch.as_tensor(D32)
    So = engine.residual(Sd, "bf16", P)
    Do = engine.residual(Dd, "bf16", P)
    # random children (some nodes childless) and winners (some out of range)
    kids = [np.sort(rng.choice(n_f, rng.integers(0, 90), replace=False)) for _ in range(n_c)]
    off = np.concatenate([[0], np.cumsum([len(k) for k in kids])]).astype(np.int32)
    ch = np.concatenate(kids).astype(np.int32)
    win = rng.integers(0, n_c, N_i).astype(np.int32)
    win[:5] = [-1, n_c, n_c + 7, -3, 2 ** 30]
    lib = _lib.load()
    nb = int(lib.mtip_group_workspace_size(N_i, n_c))
    ws = torch.empty(nb, dtype=torch.uint8, device=dev)
    w = torch.as_tensor(win).to(dev)
    _lib.check(lib.mtip_group_by_winner(_ptr(w), 1, 1, N_i, n_c, _ptr(ws), nb, engine._stream()),
               "group")
    idx = torch.full((N_i, K), -7, dtype=torch.int32, device=dev)
    val = torch.full((N_i, K), 0.0, dtype=torch.float32, device=dev)
    offd, chd = torch.as_tensor(off).to(dev), torch.as_tensor(ch).to(dev)
    _lib.check(lib.mtip_group_topk(_ptr(ws), 1, N_i, n_c, _ptr(offd), _ptr(chd), _ptr(So.data), ld,
                                   _ptr(So.sumsq), _ptr(Do.data), ld, ld, K, _ptr(idx), _ptr(val),
                                   0, engine._stream()), "group_topk")
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.check(lib.mtip_group_bad_count(_ptr(ws), 1, N_i, n_c, _ptr(bad), engine._stream()), "bad")
    torch.cuda.synchronize()
    assert int(bad.item()) == 5
    # the entry permutation: every image once, grouped by winner, left-out ones last
    po = int(lib.mtip_group_perm_offset(N_i, n_c))
    perm = ws[po:po + 4 * N_i].view(torch.int32).cpu().numpy()
    assert sorted(perm.tolist()) == list(range(N_i))
    g_of = win[perm[:N_i - 5]]
    assert np.all(np.diff(g_of) >= 0)
    assert set(perm[N_i - 5:].tolist()) == set(range(5))
    idx, val = idx.cpu().numpy(), val.cpu().numpy()
    Sb, Db = _bf16_round(S32), _bf16_round(D32)
    ss = So.sumsq.cpu().numpy()
    np.testing.assert_allclose(ss, (Sb ** 2).sum(1), rtol=1e-12)
    checked = 0
    for i in range(N_i):
        if not (0 <= win[i] < n_c):
            assert (idx[i] == -7).all()          # left out, not written
            continue
        c = kids[win[i]]
        v = ss[c] - 2.0 * (Sb[c] @ Db[i])
        order = np.lexsort((c, v))[:K]
        want_idx = np.full(K, -1)
        want_idx[:len(order)] = c[order]
        n = len(order)
        if n:
            np.testing.assert_allclose(val[i, :n], v[order], rtol=2e-5, atol=2e-4 * P)
            gaps = np.diff(np.sort(v)[:K + 1]) if len(v) > K else np.diff(np.sort(v))
            if len(gaps) == 0 or gaps.min() > 1e-4 * np.abs(v).max():
                np.testing.assert_array_equal(idx[i, :n], want_idx[:n])
                checked += 1
        assert (idx[i, n:] == -1).all() and np.isinf(val[i, n:]).all()
    assert checked > N_i // 2


def test_group_topk_e4m3_equals_brute_force():
    """mtip_group_topk_prec on e4m3 rows with per-row scales against FP64 sums
    of the dequantised values; the BF16 entry point and the precision one
    agree bitwise on bf16 operands."""
    rng = np.random.default_rng(1)
    dev = torch.device("cuda", 0)
    n_c, n_f, N_i, P, K = 19, 600, 257, 448, 8
    S32 = rng.standard_normal((n_f, P)).astype(np.float32)
    D32 = rng.standard_normal((N_i, P)).astype(np.float32) * 3.0
    ld = engine.padded(P)
    Sd = torch.zeros((n_f, ld), dtype=torch.float32, device=dev)
    Sd[:, :P] = torch.as_tensor(S32)
    Dd = torch.zeros((N_i, ld), dtype=torch.float32, device=dev)
    Dd[:, :P] = torch.as_tensor(D32)
    So, Do = engine.residual(Sd, "fp8", P), engine.residual(Dd, "fp8", P)
    Sb, Db = engine.residual(Sd, "bf16", P), engine.residual(Dd, "bf16", P)
    kids = [np.sort(rng.choice(n_f, rng.integers(0, 150), replace=False)) for _ in range(n_c)]
    off = np.concatenate([[0], np.cumsum([len(k) for k in kids])]).astype(np.int32)
    ch = np.concatenate(kids).astype(np.int32)
    win = rng.integers(0, n_c, N_i).astype(np.int32)
    lib = _lib.load()
    nb = int(lib.mtip_group_workspace_size(N_i, n_c))
    ws = torch.empty(nb, dtype=torch.uint8, device=dev)
    st = engine._stream()
    _lib.check(lib.mtip_group_by_winner(_ptr(torch.as_tensor(win).to(dev)), 1, 1, N_i, n_c,
                                        _ptr(ws), nb, st), "group")
    offd, chd = torch.as_tensor(off).to(dev), torch.as_tensor(ch).to(dev)
    out = {}
    for name, prec, S, D in (("fp8", _lib.MTIP_FP8, So, Do), ("bf16", _lib.MTIP_BF16, Sb, Db)):
        idx = torch.full((N_i, K), -7, dtype=torch.int32, device=dev)
        val = torch.zeros((N_i, K), dtype=torch.float32, device=dev)
        _lib.check(lib.mtip_group_topk_prec(prec, _ptr(ws), 1, N_i, n_c, _ptr(offd), _ptr(chd),
                                            _ptr(S.data), ld, _ptr(S.scale), _ptr(S.sumsq),
                                            _ptr(D.data), ld, _ptr(D.scale), ld, K, _ptr(idx),
                                            _ptr(val), 0, st), "group_topk_prec")
        out[name] = (idx, val)
    idx_b = torch.full((N_i, K), -7, dtype=torch.int32, device=dev)
    val_b = torch.zeros((N_i, K), dtype=torch.float32, device=dev)
    _lib.check(lib.mtip_group_topk(_ptr(ws), 1, N_i, n_c, _ptr(offd), _ptr(chd), _ptr(Sb.data), ld,
                                   _ptr(Sb.sumsq), _ptr(Db.data), ld, ld, K, _ptr(idx_b),
                                   _ptr(val_b), 0, st), "group_topk")
    torch.cuda.synchronize()
    assert torch.equal(idx_b, out["bf16"][0]) and torch.equal(val_b, out["bf16"][1])
    # FP64 brute force on the dequantised e4m3 values
    sdq = (So.data[:, :P].view(torch.float8_e4m3fn).double() * So.scale.double()[:, None]).cpu().numpy()
    ddq = (Do.data[:, :P].view(torch.float8_e4m3fn).double() * Do.scale.double()[:, None]).cpu().numpy()
    ss = So.sumsq.cpu().numpy()
    np.testing.assert_allclose(ss, (sdq ** 2).sum(1), rtol=1e-12)
    idx, val = (t.cpu().numpy() for t in out["fp8"])
    checked = 0
    for i in range(N_i):
        c = kids[win[i]]
        v = ss[c] - 2.0 * (sdq[c] @ ddq[i])
        order = np.lexsort((c, v))[:K]
        n = len(order)
        if n:
            np.testing.assert_allclose(val[i, :n], v[order], rtol=2e-5, atol=1e-4 * P)
            gaps = np.diff(np.sort(v)[:K + 1]) if len(v) > K else np.diff(np.sort(v))
            if len(gaps) == 0 or gaps.min() > 1e-4 * np.abs(v).max():
                np.testing.assert_array_equal(idx[i, :n], c[order])
                checked += 1
        assert (idx[i, n:] == -1).all() and np.isinf(val[i, n:]).all()
    assert checked > N_i // 2


def test_beam_grouping_and_distinct_merge_equal_brute_force():
    rng = np.random.default_rng(4)
    dev = torch.device("cuda", 0)
    n_c, n_f, N_i, P, T, K = 31, 500, 211, 192, 3, 4
    S32 = rng.standard_normal((n_f, P)).astype(np.float32)
    D32 = rng.standard_normal((N_i, P)).astype(np.float32)
    ld = engine.padded(P)
    Sd = torch.zeros((n_f, ld), dtype=torch.float32, device=dev)
    Sd[:, :P] = torch.as_tensor(S32)
    Dd = torch.zeros((N_i, ld), dtype=torch.float32, device=dev)
    Dd[:, :P] = torch.as_tensor(D32)
    So, Do = engine.residual(Sd, "bf16", P), engine.residual(Dd, "bf16", P)
    kids = [np.sort(rng.choice(n_f, rng.integers(1, 70), replace=False)) for _ in range(n_c)]
    off = np.concatenate([[0], np.cumsum([len(k) for k in kids])]).astype(np.int32)
    ch = np.concatenate(kids).astype(np.int32)
    # winners (N_i, 4): the first T used; repeats within a row allowed
    win = rng.integers(0, n_c, (N_i, 4)).astype(np.int32)
    win[::7, 1] = win[::7, 0]
    lib = _lib.load()
    nb = int(lib.mtip_group_workspace_size(N_i * T, n_c))
    ws = torch.empty(nb, dtype=torch.uint8, device=dev)
    w = torch.as_tensor(win).to(dev)
    _lib.check(lib.mtip_group_by_winner(_ptr(w), 4, T, N_i, n_c, _ptr(ws), nb, engine._stream()),
               "group")
    eidx = torch.full((N_i * T, K), -1, dtype=torch.int32, device=dev)
    evals = torch.full((N_i * T, K), float("inf"), dtype=torch.float32, device=dev)
    offd, chd = torch.as_tensor(off).to(dev), torch.as_tensor(ch).to(dev)
    _lib.check(lib.mtip_group_topk(_ptr(ws), T, N_i, n_c, _ptr(offd), _ptr(chd), _ptr(So.data),
                                   ld, _ptr(So.sumsq), _ptr(Do.data), ld, ld, K, _ptr(eidx),
                                   _ptr(evals), 0, engine._stream()), "group_topk")
    idx = torch.empty((N_i, K), dtype=torch.int32, device=dev)
    val = torch.empty((N_i, K), dtype=torch.float32, device=dev)
    _lib.check(lib.mtip_merge_distinct(_ptr(eidx), _ptr(evals), N_i, T * K, K, _ptr(idx),
                                       _ptr(val), engine._stream()), "merge")
    torch.cuda.synchronize()
    idx, val = idx.cpu().numpy(), val.cpu().numpy()
    Sb, Db = _bf16_round(S32), _bf16_round(D32)
    ss = So.sumsq.cpu().numpy()
    checked = 0
    for i in range(N_i):
        c = np.unique(np.concatenate([kids[g] for g in win[i, :T]]))
        v = ss[c] - 2.0 * (Sb[c] @ Db[i])
        order = np.lexsort((c, v))[:K]
        n = len(order)
        assert len(set(idx[i, :n].tolist())) == n          # distinct
        np.testing.assert_allclose(val[i, :n], v[order], rtol=2e-5, atol=2e-4 * P)
        srt = np.sort(v)[:K + 1]
        if len(srt) < 2 or np.diff(srt).min() > 1e-4 * np.abs(v).max():
            np.testing.assert_array_equal(idx[i, :n], c[order])
            checked += 1
    assert checked > N_i // 2


@pytest.fixture(scope="module")
def small_hier():
    M, atoms = 33, 150
    ac = synth.make_model(M, atoms, 3)
    ppr, pdr = synth.detector(1, 32, 32)
    E = float(pdr.max())
    h = hierarchy.Hierarchy(sizes=(300, 3000, 12000), n_parents=8)
    plan = engine.NufftPlan(M, 1e-6, 0)
    plan.set_model(ac)
    q = np.ascontiguousarray(ppr.reshape(3, -1))
    hm = hierarchy.HierarchyModel(h, plan, q, np.pi / E, order=engine.tiling(ppr.shape[1:]))
    torch.cuda.synchronize()
    assert hm.bad == 0
    return dict(ac=ac, ppr=ppr, pdr=pdr, h=h, hm=hm, plan=plan, E=E)


def _planted(hm, n, rng, scale=2.5):
    """Images equal to scaled last-level slices (an exact answer exists)."""
    last = hm.levels[-1]
    pick = rng.choice(last.N_o, n, replace=False)
    imgs = (scale * last.f32[torch.as_tensor(pick).to(last.f32.device), :last.N_p]).contiguous()
    return pick, imgs


def test_hierarchy_finds_planted_and_agrees_with_flat(small_hier):
    hm = small_hier["hm"]
    rng = np.random.default_rng(1)
    pick, imgs = _planted(hm, 400, rng, scale=1.0)
    sizes = np.array([7, 150, 33, 210])
    offs = np.concatenate([[0], np.cumsum(sizes)])
    scales = np.repeat([1.0, 3.0, 0.5, 2.0], sizes)
    imgs = imgs * torch.as_tensor(scales, dtype=torch.float32, device=imgs.device)[:, None]
    iset = engine.ImageSet(imgs)
    # the exact planting scale, so only the search is under test
    alpha = torch.as_tensor(scales, dtype=torch.float64).to(imgs.device)
    hmatch = hierarchy.HierarchicalMatcher(hm, K=8)
    r = hmatch.run(iset, offs, alpha=alpha)
    got = r["index"].cpu().numpy()
    flat = engine.match_resident(hm.levels[-1], iset, "bf16", K=8, alpha=alpha)
    flat_idx = flat["index"].cpu().numpy()
    assert (flat_idx == pick).mean() >= 0.99
    assert (got == pick).mean() >= 0.93
    assert (got == flat_idx).mean() >= 0.93
    # the per-batch std ratio: one value per batch, close to the batch's scale
    a = hmatch.run(iset, offs)["alpha"].cpu().numpy()
    for b0, b1, sc in zip(offs[:-1], offs[1:], [1.0, 3.0, 0.5, 2.0]):
        assert np.all(a[b0:b1] == a[b0]) and abs(a[b0] / sc - 1.0) < 0.3


@pytest.mark.parametrize("fp8_levels", [(), (1,)])
def test_hierarchy_noisy_images_agree_with_flat(small_hier, fp8_levels):
    hm, ac, ppr, E = small_hier["hm"], small_hier["ac"], small_hier["ppr"], small_hier["E"]
    rng = np.random.default_rng(2)
    true = synth.random_quaternions(300, rng)
    sl = oracle_slicer(1e-9)(ac, true, ppr, E)
    imgs = synth.poisson_images(sl, 5.0, rng)
    iset = engine.ImageSet(imgs.reshape(len(imgs), -1))
    offs = [0, 100, 300]
    r = hierarchy.HierarchicalMatcher(hm, K=8, fp8_levels=fp8_levels).run(iset, offs)
    flat = engine.match_resident(hm.levels[-1], iset, "bf16", K=8, alpha=r["alpha"])
    agree = (r["index"].cpu().numpy() == flat["index"].cpu().numpy()).mean()
    assert agree >= 0.85


@pytest.mark.parametrize("fp8_levels", [(), (1,)])
def test_graph_replay_equals_eager(small_hier, fp8_levels):
    hm = small_hier["hm"]
    rng = np.random.default_rng(3)
    _, imgs = _planted(hm, 257, rng)
    imgs = imgs + torch.rand_like(imgs)
    iset = engine.ImageSet(imgs)
    offs = [0, 17, 100, 257]
    m = hierarchy.HierarchicalMatcher(hm, K=8, fp8_levels=fp8_levels)
    eager = m.run(iset, offs)
    e_idx, e_d = eager["index"].cpu().numpy(), eager["best_d64"].cpu().numpy()
    m.capture(iset, offs)
    for rep in range(3):
        out = m.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out["index"].cpu().numpy(), e_idx)
        np.testing.assert_array_equal(out["best_d64"].cpu().numpy(), e_d)
        # unrelated eager work between replays
        again = m.run(iset, offs)
        np.testing.assert_array_equal(again["index"].cpu().numpy(), e_idx)


@pytest.mark.parametrize("M,atoms", [(81, 800), (121, 2700)])
def test_config3_hierarchy_agrees_with_flat_and_replays(M, atoms):
    """The batching form at its configured size: 4 x 32 x 128 px, the 1k /
    10k / 100k hierarchy, uneven Poisson batches (sizes log-uniform in
    [16, 512], 2,000+ images in all); the hierarchical result equals the flat
    BF16 + FP64 search over the 100k level for at least 99 % of the images, and
    the captured graph's replays equal the eager pass bitwise."""
    ppr, pdr = synth.detector(4, 32, 128)
    E = float(pdr.max())
    print(f"case: config-3 hierarchy M={M}", flush=True)
    h = hierarchy.Hierarchy(sizes=(1000, 10000, 100000), n_parents=4)
    ac = synth.make_model(M, atoms, 0)
    plan = engine.NufftPlan(M, 1e-6, 0)
    plan.set_model(ac)
    q = torch.as_tensor(np.ascontiguousarray(ppr.reshape(3, -1))).cuda()
    tiling = engine.auto_tiling(plan, ppr.shape[1:])
    hm = hierarchy.HierarchyModel(h, plan, q, np.pi / E, order=tiling)
    sizes = hierarchy.log_uniform_batches(12, 16, 512, synth.rng_for(M, 7))
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    N_i = int(offs[-1])
    true = synth.random_quaternions(N_i, synth.rng_for(M, 8))
    sl = plan.slices(torch.as_tensor(true).cuda(), q, np.pi / E, order=tiling)["s32"][:, :q.shape[1]]
    sl.clamp_(min=0.0)
    sl.mul_(1.0 / float(sl.mean()))
    g = torch.Generator(device="cuda")
    g.manual_seed(M)
    iset = engine.ImageSet(torch.poisson(sl, generator=g))
    m = hierarchy.HierarchicalMatcher(hm, K=8)
    eager = m.run(iset, offs)
    e_idx, e_d = eager["index"].cpu().numpy(), eager["best_d64"].cpu().numpy()
    flat = engine.match_resident(hm.levels[-1], iset, "bf16", K=8, alpha=eager["alpha"])
    agree = float((e_idx == flat["index"].cpu().numpy()).mean())
    print(f"config-3 M={M}: {N_i} images, agreement with the flat 100k search {agree:.4f}",
          flush=True)
    assert agree >= 0.99
    m.capture(iset, offs)
    for rep in range(3):
        out = m.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out["index"].cpu().numpy(), e_idx)
        np.testing.assert_array_equal(out["best_d64"].cpu().numpy(), e_d)
        again = m.run(iset, offs)
        np.testing.assert_array_equal(again["index"].cpu().numpy(), e_idx)
