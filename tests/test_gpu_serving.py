"""Serving form on the GPU: continuous batching gives each request the same
result alone and in the mix, and matches the oracle."""
import numpy as np
import pytest

from helpers import oracle_slicer
from oracle import matching
from spinifel_amd import synth

pytestmark = pytest.mark.gpu


def _model(M, atoms, seed, n_ref, panels=1, rows=32, cols=32):
    ac = synth.make_model(M, atoms, seed)
    ppr, pdr = synth.detector(panels, rows, cols)
    ref = synth.random_quaternions(n_ref, np.random.default_rng(seed + 100))
    return ac, ppr, pdr, ref


@pytest.fixture(scope="module")
def server():
    from spinifel_amd import serving
    srv = serving.Server(precision="bf16", K=8, n_slots=3, max_batch_images=600)
    models = {}
    for name, (M, atoms, seed, n_ref) in {"a": (33, 120, 1, 1200), "b": (41, 300, 2, 2000)}.items():
        ac, ppr, pdr, ref = _model(M, atoms, seed, n_ref)
        srv.add_model(name, ac, ref, ppr, pdr)
        true = synth.random_quaternions(400, np.random.default_rng(seed + 7))
        sl = oracle_slicer(1e-9)(ac, true, ppr, pdr.max())
        imgs = synth.poisson_images(sl, 0.5, np.random.default_rng(seed + 9))
        models[name] = dict(ac=ac, ppr=ppr, pdr=pdr, ref=ref, imgs=imgs)
    return srv, models


def test_requests_alone_equal_in_mix(server):
    srv, models = server
    rng = np.random.default_rng(0)
    reqs = []
    for _ in range(24):
        name = "ab"[int(rng.integers(0, 2))]
        n = int(rng.integers(1, 90))
        sel = rng.choice(400, n, replace=False)
        reqs.append((name, models[name]["imgs"][sel]))
    alone = [srv.match_now(name, imgs) for name, imgs in reqs]
    srv.start()
    futs = [srv.submit(name, imgs) for name, imgs in reqs]
    mixed = [f.result(timeout=120) for f in futs]
    srv.stop()
    batched = sum(1 for r in mixed if r.batch_images > 0)
    assert batched == len(reqs)
    assert max(r.batch_images for r in mixed) > max(len(i) for _, i in reqs), "no batching happened"
    for a, m in zip(alone, mixed):
        np.testing.assert_array_equal(a.index, m.index)
        np.testing.assert_array_equal(a.dist64, m.dist64)     # bitwise: per-image arithmetic only
    st = srv.stats()
    assert st["requests"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0


def test_pinned_tensor_requests_equal_numpy(server):
    """Requests handed in as pinned host tensors (uploaded without a host copy)
    batch with numpy requests and give bitwise the numpy results; a tensor that
    is not pinned host float32 is refused."""
    import torch
    srv, models = server
    rng = np.random.default_rng(5)
    reqs = []
    for k in range(16):
        name = "ab"[int(rng.integers(0, 2))]
        sel = rng.choice(400, int(rng.integers(1, 120)), replace=False)
        reqs.append((name, models[name]["imgs"][sel].astype(np.float32), k % 2 == 0))
    alone = [srv.match_now(name, imgs) for name, imgs, _ in reqs]
    srv.start()
    try:
        futs = [srv.submit(name, torch.from_numpy(imgs.reshape(len(imgs), -1)).pin_memory()
                           if pinned else imgs)
                for name, imgs, pinned in reqs]
        mixed = [f.result(timeout=120) for f in futs]
        with pytest.raises(ValueError):
            srv.submit("a", torch.from_numpy(reqs[0][1].reshape(len(reqs[0][1]), -1)))
    finally:
        srv.stop()
    assert max(r.batch_images for r in mixed) > max(len(i) for _, i, _ in reqs), "no batching"
    for a, m in zip(alone, mixed):
        np.testing.assert_array_equal(a.index, m.index)
        np.testing.assert_array_equal(a.dist64, m.dist64)


def test_served_result_matches_oracle(server):
    srv, models = server
    m = models["b"]
    imgs = m["imgs"][:200]
    got = srv.match_now("b", imgs)
    ora = matching.slicing_and_match(m["ac"], imgs.reshape((200, 1, 32, 32)), m["ppr"], m["pdr"],
                                     m["ref"], return_details=True)
    assert (got.index == ora["index"]).mean() == 1.0
    # FP64 distance of the winner equals the oracle's squared distance
    np.testing.assert_allclose(got.dist64, ora["min_dist"] ** 2, rtol=1e-5)


def test_admission_waits_for_memory(server):
    """With the reserve set above the free HBM nothing is admitted; lowering it
    releases the queue (the host-side admission rule, forced from the host)."""
    from spinifel_amd import serving
    srv, models = server
    free, _ = serving.free_hbm(0)
    old = srv.mem_reserve
    srv.mem_reserve = free + (1 << 40)
    srv.start()
    fut = srv.submit("a", models["a"]["imgs"][:10])
    import time
    time.sleep(0.3)
    assert not fut.done()
    srv.mem_reserve = old
    res = fut.result(timeout=60)
    srv.stop()
    assert len(res.index) == 10


def test_admission_gives_up_when_nothing_can_free_memory(server):
    """A batch that cannot fit while every slot is idle fails with MemoryError
    after the admission timeout (released cache included) instead of waiting
    forever; the server keeps serving afterwards."""
    from spinifel_amd import serving
    srv, models = server
    free, _ = serving.free_hbm(0)
    old_res, old_to = srv.mem_reserve, srv.admit_timeout_s
    srv.mem_reserve = free + (1 << 40)
    srv.admit_timeout_s = 0.2
    srv.start()
    try:
        fut = srv.submit("a", models["a"]["imgs"][:10])
        with pytest.raises(MemoryError):
            fut.result(timeout=30)
        srv.mem_reserve = old_res
        res = srv.submit("a", models["a"]["imgs"][:10]).result(timeout=60)
        assert len(res.index) == 10
    finally:
        srv.mem_reserve, srv.admit_timeout_s = old_res, old_to
        srv.stop()


def test_config4_models_alone_equal_in_mix():
    """The serving form at its configured scale: four resident models (M = 81,
    81, 121, 121 with 10,000 / 100,000 / 10,000 / 100,000 orientations, 4 x 32 x
    128 px, bf16 reference slices); a sample of Poisson-burst requests (sizes
    in [1, 512]) gives bitwise the same orientations and FP64 distances alone
    and in the mix."""
    import torch
    from spinifel_amd import engine, serving
    ppr, pdr = synth.detector(4, 32, 128)
    E = float(pdr.max())
    srv = serving.Server(precision="bf16", K=8, n_slots=4, max_batch_images=4096)
    spec = [("m81_10k", 81, 800, 10000), ("m81_100k", 81, 800, 100000),
            ("m121_10k", 121, 2700, 10000), ("m121_100k", 121, 2700, 100000)]
    pools = {}
    for k, (name, M, atoms, n_ref) in enumerate(spec):
        print(f"serving config 4: model {name}", flush=True)
        ac = synth.make_model(M, atoms, k)
        ref = synth.random_quaternions(n_ref, synth.rng_for(k, 9))
        srv.add_model(name, ac, ref, ppr, pdr)
        m = srv.models[name]
        assert m.tile == 256
        true = synth.random_quaternions(600, synth.rng_for(k, 10))
        plan = srv._plans[(M, srv.eps)]
        q = torch.as_tensor(np.ascontiguousarray(ppr.reshape(3, -1))).to(srv.device)
        sl = plan.slices(torch.as_tensor(true).to(srv.device), q, np.pi / E,
                         order=engine.auto_tiling(plan, ppr.shape[1:]))["s32"][:, :q.shape[1]]
        imgs = synth.poisson_images(sl.double().cpu().numpy(), 1.0, synth.rng_for(k, 11))
        pools[name] = imgs
    rng = synth.rng_for(4, 4)
    sched = serving.poisson_bursts(120.0, 0.2, 512, [s[0] for s in spec], rng)
    reqs = [(name, pools[name][rng.integers(0, 600, n)]) for _, name, n in sched[:16]]
    print(f"serving config 4: {len(reqs)} requests, {sum(len(i) for _, i in reqs)} images",
          flush=True)
    alone = [srv.match_now(name, imgs) for name, imgs in reqs]
    srv.start()
    futs = [srv.submit(name, imgs) for name, imgs in reqs]
    mixed = [f.result(timeout=300) for f in futs]
    srv.stop()
    assert max(r.batch_images for r in mixed) > max(len(i) for _, i in reqs), "no batching happened"
    for a, m in zip(alone, mixed):
        np.testing.assert_array_equal(a.index, m.index)
        np.testing.assert_array_equal(a.dist64, m.dist64)
