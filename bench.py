#!/usr/bin/env python3
"""Orientation-matching benchmark on one MI355X.

One step = one ``match`` of N_i resident images against N_o reference
orientations of a resident model: fine-grid load of the model, FP64
rotation + ES-kernel NUFFT slicing of all orientations, std scaling,
centred residual operands, MFMA cross term with fused per-image top-K,
merge, and (BF16 / FP8 forms) FP32 + FP64 rescoring of 8 candidates.
Image statistics are computed once when the images become resident (the
reference's class path does the same, spinifel/sequential/orientation_matching.py:70-76).

Default workload: M=81 model (800 Gaussian atoms), 4 x 32 x 128 = 16,384
Ewald-sphere pixels, 10,000 random reference quaternions, 10,000 Poisson
images (1 photon/px) at random true orientations; synthetic, seeded.

Prints one JSON line.  ``value`` = images/s of the BF16 + rescore form;
the FP32 (plain) and FP8 forms are timed the same way and reported under
``forms``.  Accuracy for every form over all images: argmin agreement and
relative distance error against FP64 distances of every image x orientation
pair (an FP64 library GEMM on the GPU as the checker, from the same
slices); agreement with the oracle's FP64 argmin (NUFFT eps 1e-12) on the
committed fixture tests/golden/m81_16k_px.npz; NUFFT slice error against
the FP64 direct NUDFT on a pixel subsample.  ``dropin_host_io`` times the
drop-in ``match`` as the reference calls it, host arrays in and out.
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"hbm_gbs": 8000.0, "fp32": 157.3, "bf16": 2500.0, "fp8": 5000.0}   # GB/s, TFLOP/s
_PREC_CODE = {"fp32": 0, "bf16": 1, "fp8": 2}


def p_n_ref(W):
    return int(W["p"]["n_ref"])


def p_n_img(W):
    return int(W["p"]["n_img"])


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="m81_16k_px")
    p.add_argument("--forms", default="fp32,bf16,fp8")
    p.add_argument("--n-ref", type=int, default=None)
    p.add_argument("--n-img", type=int, default=None)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--eps", type=float, default=1e-6)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-check", action="store_true", help="skip the accuracy checks")
    p.add_argument("--cpu-seconds", type=float, default=20.0,
                   help="target CPU time of the baseline sample")
    p.add_argument("--no-tune", action="store_true", help="rule-based tiles, no autotuning")
    p.add_argument("--no-batching", action="store_true", help="skip the batching form")
    p.add_argument("--no-serving", action="store_true", help="skip the serving form")
    p.add_argument("--no-dropin", action="store_true",
                   help="skip timing the drop-in match() with host inputs")
    p.add_argument("--no-merge", action="store_true",
                   help="skip the merge (type-1 + CG) and phasing leg")
    p.add_argument("--merge-only", action="store_true",
                   help="run only the merge leg (for profiling its kernels)")
    p.add_argument("--merge-images", type=int, default=1000,
                   help="images of the merge leg (x 16,384 px)")
    p.add_argument("--batch-steps", type=int, default=5, help="timed graph replays per model")
    p.add_argument("--serve-seconds", type=float, default=8.0,
                   help="length of the Poisson arrival schedule")
    p.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                   help="only copy BYTES device-to-device 3 times (hardware-counter "
                        "calibration of a plain copy of known size) and exit")
    return p.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(args, torch, engine, synth):
    p = dict(synth.WORKLOADS[args.workload])
    if args.n_ref:
        p["n_ref"] = args.n_ref
    if args.n_img:
        p["n_img"] = args.n_img
    dev = torch.device("cuda", 0)
    ac = synth.make_model(p["M"], p["atoms"], args.seed)
    ppr, pdr = synth.detector(p["panels"], p["rows"], p["cols"])
    E = float(pdr.max())
    ref = synth.random_quaternions(p["n_ref"], synth.rng_for(args.seed, synth.STREAM_REF))
    true = synth.random_quaternions(p["n_img"], synth.rng_for(args.seed, synth.STREAM_TRUE))
    plan = engine.NufftPlan(p["M"], args.eps, 0)
    plan.set_model(ac)
    q = torch.as_tensor(np.ascontiguousarray(ppr.reshape(3, -1))).to(dev)
    scale = np.pi / E
    # pixel tiling of the NUFFT: autotuned (or by rule)
    if args.no_tune:
        order, tune = engine.auto_tiling(plan, ppr.shape[1:]), None
    else:
        from spinifel_amd import autotune
        order, tune = autotune.nufft_tiling(plan, q, np.pi / E, ppr.shape[1:])
    # images: GPU slices at the true orientations, clipped, scaled, Poisson (numpy, seeded)
    true_sl = plan.slices(torch.as_tensor(true).to(dev), q, scale, order=order)["s32"][:, :q.shape[1]]
    imgs = synth.poisson_images(true_sl.double().cpu().numpy(), p["mean_photons"],
                                synth.rng_for(args.seed, synth.STREAM_POISSON))
    del true_sl
    ac_dev = torch.as_tensor(ac).to(dev)
    quats_dev = torch.as_tensor(ref).to(dev)
    return dict(p=p, ac=ac, ac_dev=ac_dev, ppr=ppr, pdr=pdr, E=E, ref=ref, true=true, plan=plan,
                q=q, scale=scale, quats=quats_dev, images=imgs, dev=dev, order=order,
                nufft_tune=dict(tile=list(order.shape), ms=tune) if order is not None else None)


def run_form(form, W, args, torch, engine, timers):
    precision = form
    K = 1 if form == "fp32" else 8
    resc = form != "fp32"
    imgs = W["imgset"]
    plan = W["plan"]

    def step(t=None):
        plan.set_model(W["ac_dev"])
        sl = engine.SliceSet.compute(plan, W["quats"], W["q"], W["scale"], timers=t, order=W["order"])
        return sl, engine.match_resident(sl, imgs, precision, K=K, rescore_fp64=resc, timers=t)

    # the interpreter's cyclic collector runs before the warm-up, not inside
    # the steps (a collection inside a step stalls the launches and the GPU
    # idles) nor between warm-up and timing (the GPU clock drops while idle)
    gc.collect()
    gc.disable()
    try:
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        timers.clear()
        lat = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            s0 = time.perf_counter()
            sl, r = step(timers)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - s0)
        total = time.perf_counter() - t0
    finally:
        gc.enable()
    N_i = imgs.N_i
    out = dict(images_per_s=N_i * args.steps / total, ms_per_step=1e3 * total / args.steps,
               p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)),
               K=K, rescore="fp32+fp64" if resc else None)
    kern = {}
    for name in list(timers.spans):
        ms = timers.ms(name)
        kern[name] = float(np.mean(ms))
    out["kernel_ms"] = kern
    # slices resident (serving / batching view): matching only
    for _ in range(2):
        engine.match_resident(sl, imgs, precision, K=K, rescore_fp64=resc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = engine.match_resident(sl, imgs, precision, K=K, rescore_fp64=resc)
    torch.cuda.synchronize()
    tr = (time.perf_counter() - t0) / args.steps
    tres = engine.Timers()
    engine.match_resident(sl, imgs, precision, K=K, rescore_fp64=resc, timers=tres)
    torch.cuda.synchronize()
    out["slices_resident"] = dict(images_per_s=N_i / tr, ms_per_match=1e3 * tr,
                                  kernel_ms={k: float(np.mean(tres.ms(k))) for k in list(tres.spans)})
    return out, sl, r


def measured_traffic(workload):
    """HBM bytes per launch of each kernel from the committed counter summary
    (profiles/traffic_<workload>.json: FETCH_SIZE + WRITE_SIZE per dispatch,
    scaled by the counters' reading for a plain copy of known size)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch", {})
    except (OSError, ValueError):
        return {}


def roofline_entries(W, kern_ms, forms, engine):
    p = W["p"]
    N_o, N_i = p["n_ref"], p["n_img"]
    N_p = int(W["q"].shape[1])
    P = N_o * N_p
    npad = W["plan"].padded_edge
    traffic = measured_traffic(W["workload"])
    ent = {}
    if "nufft" in kern_ms:
        # FP32 slice write + real fine grid read once + FP64 q + FP64 quats + FP64 row sums
        byts = 4 * P + 4 * npad ** 3 + 24 * N_p + 32 * N_o + 16 * N_o
        t = kern_ms["nufft"] * 1e-3
        ent["nufft"] = dict(bound="hbm", achieved=byts / t / 1e9, peak=PEAK["hbm_gbs"], unit="GB/s",
                            frac=byts / t / 1e9 / PEAK["hbm_gbs"], traffic=traffic.get("nufft"),
                            algorithmic_bytes=byts, ms=kern_ms["nufft"],
                            taps_per_s=P * W["plan"].width ** 3 / t)
    for f in forms:
        key = "gemm_" + f
        if key in kern_ms:
            flops = 2.0 * N_o * N_i * N_p
            t = kern_ms[key] * 1e-3
            ent[key] = dict(bound="mfma", achieved=flops / t / 1e12, peak=PEAK[f], unit="TFLOP/s",
                            frac=flops / t / 1e12 / PEAK[f], traffic=traffic.get(key),
                            algorithmic_flops=flops,
                            ms=kern_ms[key])
    return ent


def accuracy(form_res, W, sl, torch, chunk=2000):
    """Argmin agreement and distance errors of every form over all images,
    against FP64 distances ||alpha s - d||^2 computed on the GPU (a library
    FP64 GEMM as the checker) from the same FP32 slices, alpha in FP64."""
    S = sl.f32[:, :sl.N_p].double()
    D_all = W["imgset"].f32[:, :sl.N_p]
    N_i = D_all.shape[0]
    n_tot = float(N_i * sl.N_p)
    alpha = float(torch.sqrt(D_all.double().pow(2).sum() / n_tot - (D_all.double().sum() / n_tot) ** 2)
                  / S.std(unbiased=False))
    s2 = (S * S).sum(1)
    ref_idx = torch.empty(N_i, dtype=torch.int64, device=S.device)
    ref_d2 = torch.empty((N_i, 2), dtype=torch.float64, device=S.device)
    picks = {f: dict(top=r["topk_idx"][:, 0].long(), idx=r["index"].long()) for f, r in form_res.items()}
    d2_at = {f: dict(top=torch.empty(N_i, dtype=torch.float64, device=S.device),
                     idx=torch.empty(N_i, dtype=torch.float64, device=S.device)) for f in form_res}
    for c0 in range(0, N_i, chunk):
        Dc = D_all[c0:c0 + chunk].double()
        d2 = (alpha * alpha) * s2[:, None] - (2 * alpha) * (S @ Dc.T) + (Dc * Dc).sum(1)[None, :]
        v, i = torch.topk(d2, 2, dim=0, largest=False, sorted=True)
        ref_idx[c0:c0 + chunk] = torch.argmin(d2, dim=0)    # first index of the minimum
        ref_d2[c0:c0 + chunk] = v.T
        cols = torch.arange(Dc.shape[0], device=S.device)
        for f in form_res:
            for k in ("top", "idx"):
                d2_at[f][k][c0:c0 + chunk] = d2[picks[f][k][c0:c0 + chunk], cols]
        del d2, Dc
    margin = (ref_d2[:, 1] - ref_d2[:, 0]) / ref_d2[:, 0]
    c = sl.center[:sl.N_p].double()
    out = {}
    for f, r in form_res.items():
        a = float(r["alpha"].item())
        agree = float((picks[f]["idx"] == ref_idx).double().mean())
        # the GEMM's own estimate of its best candidate's distance:
        # a^2 (v + ||d / a - c||^2) with v = ||s'||^2 - 2 s'.d' on the centred operands
        v = r["topk_val"][:, 0].double()
        dn = torch.empty(N_i, dtype=torch.float64, device=S.device)
        for c0 in range(0, N_i, chunk):
            Dc = D_all[c0:c0 + chunk].double()
            dn[c0:c0 + chunk] = ((Dc / a - c[None, :]) ** 2).sum(1)
        est = a * a * (v + dn)
        e_gemm = (est - d2_at[f]["top"]).abs() / d2_at[f]["top"]
        ent = dict(n_images=int(N_i), argmin_agreement=agree,
                   n_disagree=int((picks[f]["idx"] != ref_idx).sum()),
                   gemm_rel_dist_err_max=float(e_gemm.max()),
                   gemm_rel_dist_err_mean=float(e_gemm.mean()))
        if "best_d64" in r:
            e = (r["best_d64"].double() - ref_d2[:, 0]).abs() / ref_d2[:, 0]
            ent.update(final_rel_dist_err_max=float(e.max()), final_rel_dist_err_mean=float(e.mean()))
        out[f] = ent
    fp32_bound = out["fp32"]["gemm_rel_dist_err_max"] if "fp32" in out else None
    out["_reference"] = dict(
        n_images=int(N_i), alpha_fp64=alpha, min_rel_margin=float(margin.min()),
        fp32_rel_dist_err_bound=fp32_bound,
        n_margin_below_fp32_bound=(int((margin < fp32_bound).sum()) if fp32_bound else None),
        n_margin_below_1e6=int((margin < 1e-6).sum()),
        checker="FP64 distances of all images x all orientations on the GPU (torch FP64 "
                "GEMM) from the same FP32 slices")
    return out


GOLDEN_M81 = os.path.join(ROOT, "tests", "golden", "m81_16k_px.npz")


def golden_agreement(forms, engine, log):
    """The drop-in match() of every form on the committed oracle fixture
    (M=81, 4x32x128 px, 10,000 orientations, 1,000 images; the oracle's FP64
    argmin at NUFFT eps 1e-12): agreement with the oracle's index."""
    if not os.path.exists(GOLDEN_M81):
        return None
    from spinifel_amd import orientation_matching, settings as st
    g = dict(np.load(GOLDEN_M81))
    M = int(g["M"])
    half = g["ac_half"]
    ac = np.concatenate([half, half[:-1][::-1]]).reshape(M, M, M)
    images = g["images"].astype(np.float32)
    s = st.SpinifelSettings(N_orientations=g["ref_quats"].shape[0],
                            N_batch_size=int(g["N_batch_size"]), reduced_det_shape=(4, 32, 128))
    out = dict(fixture="tests/golden/m81_16k_px.npz", n_images=int(images.shape[0]),
               orientations=int(g["ref_quats"].shape[0]))
    for f in forms:
        # the FP32 entry is the plain form (use_single_prec: K = 1, no rescoring)
        s.use_single_prec = f == "fp32"
        r = orientation_matching.slicing_and_match(ac, images, g["ppr"], g["pdr"], g["ref_quats"],
                                                   settings=s, precision=f, return_details=True)
        out[f] = dict(argmin_agreement=float((r["index"] == g["index"]).mean()),
                      alpha_rel_err=float(abs(r["alpha"][0] - g["alpha"]) / g["alpha"]))
        log("golden", f, json.dumps(out[f]))
    return out


def dropin_timing(W, forms, args, torch, log):
    """The drop-in match() as the reference calls it: host model, host float32
    images, host quaternions in, host quaternions out (every upload, the
    slicing, the matching and the download inside the timed call), with the
    default settings (use_single_prec False: the FP32 form rescores its top-4
    in FP64, the reference's float64 distance semantics)."""
    from spinifel_amd import orientation_matching, settings as st
    p = W["p"]
    det = (p["panels"], p["rows"], p["cols"])
    imgs = np.ascontiguousarray(W["images"].reshape((-1,) + det), dtype=np.float32)
    s = st.SpinifelSettings(N_orientations=p["n_ref"], N_batch_size=1000, reduced_det_shape=det,
                            nufft_eps=args.eps)
    out = dict(images_host_bytes=int(imgs.nbytes))
    for f in forms:
        def call():
            return orientation_matching.slicing_and_match(W["ac"], imgs, W["ppr"], W["pdr"],
                                                          W["ref"], settings=s, precision=f)
        call()
        torch.cuda.synchronize()
        ts = []
        for _ in range(max(3, args.steps // 4)):
            t0 = time.perf_counter()
            q = call()
            ts.append(time.perf_counter() - t0)
        assert q.shape == (imgs.shape[0], 4)
        t = float(np.median(ts))
        out[f] = dict(images_per_s=imgs.shape[0] / t, ms_per_call=1e3 * t, calls=len(ts))
        log("drop-in", f, json.dumps(out[f]))
    return out


def nufft_check(W, sl):
    from oracle import es_nufft, geometry
    rng = np.random.default_rng(7)
    o = rng.integers(0, sl.N_o, 200)
    j = rng.integers(0, sl.N_p, 200)
    H, K, L = geometry.gen_nonuniform_normalized_positions(W["ref"][o], W["ppr"], W["E"], 1)
    Np = sl.N_p
    Hs, Ks, Ls = (H.reshape(200, Np)[np.arange(200), j], K.reshape(200, Np)[np.arange(200), j],
                  L.reshape(200, Np)[np.arange(200), j])
    ref = es_nufft.nudft3d2(Hs, Ks, Ls, W["ac"].astype(np.float64)).real
    got = sl.f32[o, j].double().cpu().numpy()
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


def cpu_baseline(W, sl, args):
    """The oracle (FP64 ES-kernel NUFFT at the reference default eps 1e-12, C/OpenMP,
    plus numpy FP64 distances + argmin) on a bounded sample, scaled to the job."""
    from oracle import es_nufft, geometry, matching
    N_o, N_i = W["p"]["n_ref"], W["p"]["n_img"]
    ppr, E = W["ppr"], W["E"]
    budget = args.cpu_seconds
    t = time.perf_counter()
    b, n, w, beta = es_nufft.fine_grid(W["ac"].astype(np.float64), -1, 1e-12)
    t_grid = time.perf_counter() - t

    def slice_time(k):
        H, K, L = geometry.gen_nonuniform_normalized_positions(W["ref"][:k], ppr, E, 1)
        t0 = time.perf_counter()
        es_nufft.interp(b, H, K, L, w, beta)
        return time.perf_counter() - t0

    t1 = slice_time(4)
    n_s = int(max(4, min(N_o, 0.5 * budget / max(t1 / 4, 1e-9))))
    ts = slice_time(n_s)
    S = sl.f32[:, :sl.N_p].double().cpu().numpy()
    D = W["images"].reshape(N_i, -1).astype(np.float64)
    t0 = time.perf_counter()
    matching.nearest_neighbor(S, D[:16])
    tm1 = time.perf_counter() - t0
    n_m = int(max(16, min(N_i, 0.5 * budget / max(tm1 / 16, 1e-9))))
    t0 = time.perf_counter()
    matching.nearest_neighbor(S, D[:n_m])
    tm = time.perf_counter() - t0
    total = t_grid + ts * N_o / n_s + tm * N_i / n_m
    cores = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", cores))
    return dict(value=N_i / total, unit="images/s", cores=min(cores, threads), kind="port",
                sample=(f"oracle FP64 ES-NUFFT eps 1e-12 (w={w}) on {n_s} of {N_o} orientations "
                        f"({ts:.1f} s) + fine grid ({t_grid:.1f} s) + numpy FP64 distances/argmin "
                        f"for {n_m} of {N_i} images vs all {N_o} ({tm:.1f} s), scaled linearly "
                        f"to the full job ({total:.0f} s)"),
                seconds_full_job=total)


def gpu_images(plan, quats, q, scale, order, mean_photons, seed, torch):
    """Poisson images (float32 device rows) at the given orientations: slices
    clipped at 0, scaled to `mean_photons` per pixel, Poisson-sampled on the GPU."""
    dev = plan.device
    sl = plan.slices(torch.as_tensor(quats).to(dev), q, scale, order=order)["s32"][:, :q.shape[1]]
    sl.clamp_(min=0.0)
    sl.mul_(mean_photons / max(float(sl.mean()), 1e-30))
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    return torch.poisson(sl, generator=g)


MODELS = {81: 800, 121: 2700}    # atoms per model size (same atom density)


def run_batching(args, torch, engine, synth, log):
    """64 batches, sizes log-uniform in [16, 4096], through a 1k / 10k / 100k
    hierarchy, one captured HIP graph per model; agreement with the flat
    search over the 100k level on an image subset.  Then the same pass with
    the 10k level on e4m3 operands (fp8_level1)."""
    from spinifel_amd import autotune, hierarchy
    from spinifel_amd.quaternions import random_quaternions
    dev = torch.device("cuda", 0)
    rng = synth.rng_for(args.seed, 7)
    sizes = hierarchy.log_uniform_batches(64, 16, 4096, rng)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    N_i = int(offs[-1])
    ppr, pdr = synth.detector(4, 32, 128)
    E = float(pdr.max())
    q = torch.as_tensor(np.ascontiguousarray(ppr.reshape(3, -1))).to(dev)
    scale = np.pi / E
    t0 = time.perf_counter()
    h = hierarchy.Hierarchy(sizes=(1000, 10000, 100000), n_parents=4)
    t_h = time.perf_counter() - t0
    out = dict(n_batches=64, batch_sizes=sizes.tolist(), images=N_i, levels=h.sizes,
               n_parents=4, level0_parts=2, level0_top=8, kept_per_entry=4, K=8,
               hierarchy_build_s=t_h, models={})
    for M, atoms in MODELS.items():
        log(f"batching: model M={M}")
        ac = synth.make_model(M, atoms, args.seed)
        plan = engine.NufftPlan(M, args.eps, 0)
        plan.set_model(ac)
        if args.no_tune:
            tiling, tune = engine.auto_tiling(plan, ppr.shape[1:]), None
        else:
            tiling, tune = autotune.nufft_tiling(plan, q, scale, ppr.shape[1:])
        t0 = time.perf_counter()
        hm = hierarchy.HierarchyModel(h, plan, q, scale, order=tiling)
        torch.cuda.synchronize()
        t_sl = time.perf_counter() - t0
        true = random_quaternions(N_i, synth.rng_for(args.seed + M, 8))
        imgs = gpu_images(plan, true, q, scale, tiling, 1.0, args.seed + M, torch)
        iset = engine.ImageSet(imgs)
        del imgs
        m = hierarchy.HierarchicalMatcher(hm, K=8)
        t = engine.Timers()
        m.run(iset, offs, timers=t)            # eager pass: per-kernel times
        torch.cuda.synchronize()
        kern = {k: float(np.mean(t.ms(k))) for k in list(t.spans)}
        m.capture(iset, offs)
        for _ in range(2):
            m.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.batch_steps):
            res = m.replay()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.batch_steps
        # flat search over the finest level on a subset, same per-batch alpha
        sub = np.sort(np.random.default_rng(11).choice(N_i, min(2048, N_i), replace=False))
        subt = torch.as_tensor(sub).to(dev)
        sset = engine.ImageSet(iset.f32[subt, :iset.N_p])
        flat = engine.match_resident(hm.levels[-1], sset, "bf16", K=8,
                                     alpha=res["alpha"][subt].contiguous())
        got = res["index"][subt].cpu().numpy()
        fl = flat["index"].cpu().numpy()
        qh = h.quats[-1]
        ang = lambda a, b: np.degrees(2 * np.arccos(np.clip(np.abs((a * b).sum(1)), 0, 1)))
        out["models"][f"M{M}"] = dict(
            images_per_s=N_i / dt, ms_per_pass=1e3 * dt, graph="hipGraph replay",
            tiling=list(tiling.shape) if tiling is not None else None, nufft_tune_ms=tune,
            slices_setup_s=t_sl, slices_resident=sum(h.sizes), kernel_ms_eager=kern,
            agreement_vs_flat_100k=float((got == fl).mean()), flat_subset=int(len(sub)),
            median_angle_to_true_deg=float(np.median(ang(qh[got], true[sub]))),
            flat_median_angle_to_true_deg=float(np.median(ang(qh[fl], true[sub]))))
        # the same pass with the 10k level's cross terms on e4m3 operands
        del m, res
        m8 = hierarchy.HierarchicalMatcher(hm, K=8, fp8_levels=(1,))
        m8.capture(iset, offs)
        for _ in range(2):
            m8.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.batch_steps):
            res8 = m8.replay()
        torch.cuda.synchronize()
        dt8 = (time.perf_counter() - t0) / args.batch_steps
        got8 = res8["index"][subt].cpu().numpy()
        out["models"][f"M{M}"]["fp8_level1"] = dict(
            images_per_s=N_i / dt8, ms_per_pass=1e3 * dt8,
            agreement_vs_flat_100k=float((got8 == fl).mean()),
            median_angle_to_true_deg=float(np.median(ang(qh[got8], true[sub]))))
        log("batching", f"M{M}", json.dumps({k: v for k, v in out["models"][f"M{M}"].items()
                                             if k not in ("kernel_ms_eager", "nufft_tune_ms")}))
        del hm, m8, res8, iset, sset, flat
        torch.cuda.empty_cache()
    return out


def run_serving(args, torch, engine, synth, log):
    """Four resident models (M = 81, 81, 121, 121 with 10k / 100k orientations),
    admitted by free HBM: Poisson bursts of U[1, 512] images at 120 bursts/s
    (light load) and at 1200 bursts/s (about 300k images/s offered), then the
    light-load bursts all at once (capacity), as numpy requests and as pinned
    host tensors."""
    from spinifel_amd import autotune, serving
    from spinifel_amd.quaternions import random_quaternions
    dev = torch.device("cuda", 0)
    ppr, pdr = synth.detector(4, 32, 128)
    E = float(pdr.max())
    q = torch.as_tensor(np.ascontiguousarray(ppr.reshape(3, -1))).to(dev)
    scale = np.pi / E
    srv = serving.Server(precision="bf16", K=8, n_slots=4, max_batch_images=4096,
                         compute_streams=2)
    pools = {}
    spec = [("m81_10k", 81, 10000), ("m81_100k", 81, 100000), ("m121_10k", 121, 10000),
            ("m121_100k", 121, 100000)]
    t0 = time.perf_counter()
    for k, (name, M, n_ref) in enumerate(spec):
        ac = synth.make_model(M, MODELS[M], args.seed + k)
        plan = engine.NufftPlan(M, args.eps, 0)
        plan.set_model(ac)
        tiling = engine.auto_tiling(plan, ppr.shape[1:]) if args.no_tune else \
            autotune.nufft_tiling(plan, q, scale, ppr.shape[1:])[0]
        ref = random_quaternions(n_ref, synth.rng_for(args.seed + k, 9))
        srv.add_model(name, ac, ref, ppr, pdr, tiling=tiling)
        true = random_quaternions(1024, synth.rng_for(args.seed + k, 10))
        imgs = gpu_images(srv._plans[(M, srv.eps)], true, q, scale, tiling, 1.0, args.seed + k,
                          torch)
        pools[name] = imgs.cpu().numpy()
        del imgs
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    pinned = {n: torch.from_numpy(p).pin_memory() for n, p in pools.items()}
    names = [s[0] for s in spec]

    def requests(sched, seed, pool):
        # each burst is a contiguous run of one model's images from a random
        # start (a view: no host copy before submission)
        rng = synth.rng_for(args.seed, seed)
        starts = [int(rng.integers(0, 1024 - n + 1)) for (_, _, n) in sched]
        reqs = [pool[m][a:a + n] for (_, m, n), a in zip(sched, starts)]
        it = iter(reqs)
        return lambda name, n: next(it)

    for n in names:                        # warm-up: one request per model
        srv.match_now(n, pools[n][:8])
    out = dict(models=[dict(name=n, M=M, orientations=r) for n, M, r in spec], setup_s=t_setup,
               server=dict(n_slots=srv.n_slots, compute_streams=len(srv._streams),
                           max_batch_images=srv.max_batch_images))
    light = serving.poisson_bursts(120.0, args.serve_seconds, 512, names,
                                   synth.rng_for(args.seed, 11))
    loaded = serving.poisson_bursts(1200.0, 2.0, 512, names, synth.rng_for(args.seed, 12))
    legs = [("poisson", light, pools, "Poisson 120 bursts/s, sizes U[1, 512], numpy"),
            ("poisson_loaded", loaded, pools, "Poisson 1200 bursts/s, sizes U[1, 512], numpy"),
            ("saturated", [(0.0, m, n) for (_, m, n) in light], pools,
             "the light-load bursts submitted at once, numpy"),
            ("saturated_pinned", [(0.0, m, n) for (_, m, n) in light], pinned,
             "the light-load bursts submitted at once, pinned host tensors")]
    for k, (leg, sched, pool, what) in enumerate(legs):
        srv.reset_stats()
        st = serving.run_load(srv, sched, requests(sched, 20 + k, pool))
        dur = max(dt for dt, _, _ in sched) if sched else 0.0
        st["offered_images_per_s"] = (sum(n for _, _, n in sched) / dur) if dur > 0 else None
        st["arrival"] = what
        out[leg] = st
        log("serving", leg, json.dumps(st))
    free, total = serving.free_hbm(0)
    out.update(seconds=args.serve_seconds, hbm_free_gb=free / 1e9, hbm_total_gb=total / 1e9)
    return out


def run_merge(args, torch, engine, synth, log):
    """The merge step after matching: type-1 spreading of 1000 images x 16,384 px
    at M = 81 (data term) and M_ups = 162 (convolution kernel, ones), then
    Merge.solve_ac with 100 CG iterations; plus the oracle's C spreading on a
    sample as the CPU baseline."""
    from spinifel_amd import _lib
    from spinifel_amd.merge import Merge
    from spinifel_amd.settings import SpinifelSettings
    from spinifel_amd.quaternions import random_quaternions
    dev = torch.device("cuda", 0)
    ppr, pdr = synth.detector(1, 128, 128)
    E = float(pdr.max())
    q = torch.as_tensor(np.ascontiguousarray(ppr.reshape(3, -1))).to(dev)
    N_p = q.shape[1]
    n_img = int(args.merge_images)
    settings = SpinifelSettings(reduced_det_shape=(1, 128, 128), nufft_eps=args.eps)
    M = settings.M
    ac = synth.make_model(M, MODELS[M], args.seed)
    plan_m = engine.NufftPlan(M, args.eps, 0)
    plan_m.set_model(ac)
    quats_np = random_quaternions(n_img, synth.rng_for(args.seed, 21))
    quats = torch.as_tensor(quats_np).to(dev)
    images = gpu_images(plan_m, quats_np, q, np.pi / E, engine.auto_tiling(plan_m, (128, 128)),
                        1.0, args.seed, torch)
    mg = Merge(settings, images, ppr, pdr)
    lib = _lib.load()
    st = engine._stream()
    P = engine._ptr
    npts = n_img * N_p
    out = dict(workload=f"M={M} (M_ups={settings.M_ups}), {n_img} images x 1x128x128 px "
                        f"= {npts / 1e6:.1f}M points, eps {args.eps}, CG maxiter "
                        f"{settings.solve_ac_maxiter}")
    # ---- spatial order of the points, spreading kernel and transform passes
    # (HIP events on the launch stream), as solve_ac runs them
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t_ord = []
    for r in range(args.warmup + args.steps):
        ev[0].record()
        (x, y, z), (xd, yd, zd, vd) = mg._points(quats)
        ev[1].record()
        torch.cuda.synchronize()
        if r >= args.warmup:
            t_ord.append(ev[0].elapsed_time(ev[1]))
    out["order_ms"] = float(np.mean(t_ord))
    log("merge order", out["order_ms"])
    for name, plan, pts, vals in (("data_M81", mg._plan, (xd, yd, zd), vd),
                                  ("ones_M162", mg._plan_ups, (x, y, z), None)):
        res_buf = torch.empty((plan.M,) * 3, dtype=torch.float64, device=dev)
        bad = torch.zeros(1, dtype=torch.int32, device=dev)
        n_sp = int(pts[0].numel())
        t_sp, t_fi = [], []
        for r in range(args.warmup + args.steps):
            _lib.check(lib.mtip_plan_spread_begin(plan._h, st))
            ev[0].record()
            _lib.check(lib.mtip_plan_spread_points(plan._h, P(pts[0]), P(pts[1]), P(pts[2]), n_sp,
                                                   P(vals), P(bad), st))
            ev[1].record()
            _lib.check(lib.mtip_plan_spread_finish(plan._h, None, 1.0 / plan.M ** 3, P(res_buf),
                                                   st))
            ev[2].record()
            torch.cuda.synchronize()
            if r >= args.warmup:
                t_sp.append(ev[0].elapsed_time(ev[1]))
                t_fi.append(ev[1].elapsed_time(ev[2]))
        sp, fi = float(np.mean(t_sp)), float(np.mean(t_fi))
        npad = plan.padded_edge
        # algorithmic HBM bytes of one spread launch: the points (24 B) and
        # values (4 B) read once and one FP64 write of the padded grid
        alg = (28 if vals is not None else 24) * n_sp + 8 * npad ** 3
        lds = 2 * 4 * n_sp * plan.width ** 3        # read + write of every tap in LDS
        out[name] = dict(spread_ms=sp, finish_ms=fi, points=n_sp, tile="128 points, Morton order",
                         points_per_s=n_sp / (sp * 1e-3),
                         hbm_alg_bytes=alg, hbm_alg_gbps=alg / (sp * 1e-3) / 1e9,
                         hbm_frac=alg / (sp * 1e-3) / 8e12,
                         lds_tap_bytes=lds, lds_tap_tbps=lds / (sp * 1e-3) / 1e12,
                         lds_frac=lds / (sp * 1e-3) / (256 * 128 * 2.4e9),
                         bad=int(bad.item()))
        log("merge", name, json.dumps(out[name]))
    del x, y, z, xd, yd, zd, vd
    # ---- whole solve_ac (positions on the fly, two type-1 transforms, CG)
    t_all, its = [], []
    for r in range(max(1, args.warmup // 2) + 2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mg.solve_ac(0, quats_np)
        torch.cuda.synchronize()
        if r >= max(1, args.warmup // 2):
            t_all.append(time.perf_counter() - t0)
            its.append(mg.last["iterations"])
    # one CG operator application (both FFT convolutions) timed alone
    x = mg.last["ac"].reshape(-1).clone()
    Wop = mg.last["W"]
    for _ in range(3):
        Wop(x)
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(10):
        Wop(x)
    ev[1].record()
    torch.cuda.synchronize()
    out["solve_ac"] = dict(ms=1e3 * float(np.mean(t_all)), iterations=int(its[-1]),
                           info=int(mg.last["info"]),
                           W_apply_ms=ev[0].elapsed_time(ev[1]) / 10)
    log("merge solve_ac", json.dumps(out["solve_ac"]))
    # ---- CPU baseline: the oracle's C spreading (OpenMP slabs) + numpy FFT
    # transform, eps as the GPU, on a sample of images
    if not args.no_cpu:
        from oracle import es_nufft, geometry
        k = 8
        H, K, L = geometry.gen_nonuniform_normalized_positions(quats_np[:k], ppr, E, 1)
        vals = images[:k].double().cpu().numpy().reshape(-1)
        t0 = time.perf_counter()
        es_nufft.adjoint(vals, H, K, L, M, True, None, args.eps)
        tc = time.perf_counter() - t0
        cores = len(os.sched_getaffinity(0))
        threads = int(os.environ.get("OMP_NUM_THREADS", cores))
        out["cpu_baseline"] = dict(
            value=k * N_p / tc, unit="points/s", cores=min(cores, threads), kind="port",
            sample=f"oracle type 1 (C spreading + numpy FFT, eps {args.eps}) of {k} images "
                   f"x {N_p} px at M={M} ({tc:.2f} s)")
        log("merge cpu", json.dumps(out["cpu_baseline"]))
    del mg, images
    # ---- phasing of the matching model's autocorrelation (default settings:
    # 5 loops of 10 ER / 5 HIO / 10 ER + shrink-wrap, then 10 ER)
    from spinifel_amd import phasing
    ev2 = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    np.random.seed(args.seed)
    phasing.phase(0, ac, settings=settings)
    torch.cuda.synchronize()
    t_ph = []
    for _ in range(2):
        t0 = time.perf_counter()
        phasing.phase(0, ac, settings=settings)
        torch.cuda.synchronize()
        t_ph.append(time.perf_counter() - t0)
    n_it = settings.N_phase_loops * (2 * settings.nER + settings.nHIO) + settings.nER
    Mt = settings.M
    rho = torch.rand((Mt,) * 3, dtype=torch.float64, device=dev)
    amp = torch.rand((Mt, Mt, Mt // 2 + 1), dtype=torch.float64, device=dev)
    ph = phasing._Phaser(amp, rho > 0.5, rho, Mt)
    for _ in range(3):
        ph.step(1, 0.9)
    ev2[0].record()
    for _ in range(20):
        ph.step(1, 0.9)
    ev2[1].record()
    torch.cuda.synchronize()
    out["phasing"] = dict(ms=1e3 * float(np.mean(t_ph)), iterations=n_it,
                          iteration_ms=ev2[0].elapsed_time(ev2[1]) / 20, M=Mt)
    if not args.no_cpu:
        # the oracle (numpy FP64 complex FFTs, as the reference's CPU path) on
        # 5 ER iterations, scaled to the full schedule
        from oracle import phasing as oph
        ac_filt_, amp_, mask = oph.amplitudes_from_ac(ac)
        sup = oph.create_support_(ac_filt_, Mt, settings.Mquat)
        r0 = sup * np.random.default_rng(args.seed).uniform(size=(Mt,) * 3)
        t0 = time.perf_counter()
        for _ in range(5):
            oph.ER(r0, amp_, mask, sup, np.inf)
        tc = (time.perf_counter() - t0) / 5
        out["phasing"]["cpu_baseline"] = dict(
            value=1e3 * tc * n_it, unit="ms per phase()", kind="port", cores=1,
            sample=f"oracle ER iteration at M={Mt} (numpy FP64 FFTs), 5 iterations "
                   f"({1e3 * tc:.1f} ms each) x {n_it}")
    log("phasing", json.dumps(out["phasing"]))
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)
    import torch
    from spinifel_amd import engine, synth
    engine.require_gpu()
    if args.copy_probe:
        n = int(args.copy_probe)
        a = torch.empty(n // 4, dtype=torch.float32, device="cuda")
        a.uniform_()
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n}), flush=True)
        return
    if args.merge_only:
        print(json.dumps({"merge": run_merge(args, torch, engine, synth, log)}), flush=True)
        return
    forms = [f for f in args.forms.split(",") if f]
    log("building workload", args.workload)
    W = build_workload(args, torch, engine, synth)
    W["workload"] = args.workload
    W["imgset"] = engine.ImageSet(W["images"].reshape(W["images"].shape[0], -1), W["dev"])
    torch.cuda.synchronize()
    gemm_tune = {}
    if not args.no_tune:
        from spinifel_amd import autotune
        log("autotuning GEMM tiles")
        sl0 = engine.SliceSet.compute(W["plan"], W["quats"], W["q"], W["scale"], order=W["order"])
        im = W["imgset"]
        a0 = engine.scaling_ratio(im.row_sum, im.row_sumsq, sl0.row_sum, sl0.row_sumsq, im.N_p)
        for f in forms:
            if f == "fp32":
                continue
            S, D = sl0.operand(f, True), im.operand(f, a0, sl0.center)
            rows, ms = autotune.gemm_tile(f, S, sl0.N_o, D, im.N_i, im.N_p, K=8)
            gemm_tune[f] = dict(rows=rows, ms=ms)
        del sl0, S, D
        torch.cuda.empty_cache()
    res, last, timers = {}, {}, {}
    sl = None
    for f in forms:
        log("timing form", f)
        t = engine.Timers()
        out, sl, r = run_form(f, W, args, torch, engine, t)
        res[f] = out
        last[f] = r
        log(f, json.dumps({k: v for k, v in out.items() if k != "kernel_ms"}), out["kernel_ms"])
    kern_all = {}
    for f in forms:
        kern_all.update(res[f]["kernel_ms"])
    roof = roofline_entries(W, kern_all, forms, engine)
    if not args.no_check:
        log("accuracy checks")
        acc = accuracy(last, W, sl, torch)
        for f in forms:
            res[f]["accuracy"] = acc[f]
        nufft_err = nufft_check(W, sl)
    else:
        acc, nufft_err = {}, None
    head = "bf16" if "bf16" in forms else forms[0]
    hk = res[head]["kernel_ms"]
    dom = max([k for k in hk if k in roof], key=lambda k: hk[k])
    rl = {k: roof[dom][k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")}
    rl["kernel"] = dom
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(W, sl, args)
    gemm_tile = {f: int(engine._lib.load().mtip_gemm_tile(_PREC_CODE[f], p_n_ref(W), p_n_img(W)))
                 for f in forms}
    del W["imgset"], sl, last
    torch.cuda.empty_cache()
    golden = None
    if not args.no_check and args.workload == "m81_16k_px":
        log("oracle fixture")
        golden = golden_agreement(forms, engine, log)
    dropin = None
    if not args.no_dropin:
        log("drop-in match() with host inputs")
        dropin = dropin_timing(W, forms, args, torch, log)
    torch.cuda.empty_cache()
    batching = None if args.no_batching else run_batching(args, torch, engine, synth, log)
    torch.cuda.empty_cache()
    serving_res = None if args.no_serving else run_serving(args, torch, engine, synth, log)
    torch.cuda.empty_cache()
    merge_res = None if args.no_merge else run_merge(args, torch, engine, synth, log)
    p = W["p"]
    line = {
        "metric": "images/s oriented, 10k orients x 16k px: BF16+rescore vs FP32; p50/p99 latency",
        "value": res[head]["images_per_s"],
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res[head]["ms_per_step"],
        "p50_ms": res[head]["p50_ms"],
        "p99_ms": res[head]["p99_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head,
        "data": "synthetic (seeded Gaussian-atom model, Poisson images)",
        "config": {"workload": (f"{args.workload}: M={p['M']} model, {p['panels']}x{p['rows']}x"
                                f"{p['cols']} px, {p['n_ref']} orientations, {p['n_img']} images, "
                                f"NUFFT eps {args.eps}"),
                   "form": f"{head}" + (" + top-8 FP32/FP64 rescore" if head != "fp32" else "")},
        "speedup_vs_fp32": (res[head]["images_per_s"] / res["fp32"]["images_per_s"]
                            if "fp32" in res and head != "fp32" else None),
        "speedup_vs_fp32_slices_resident": (
            res[head]["slices_resident"]["images_per_s"] / res["fp32"]["slices_resident"]["images_per_s"]
            if "fp32" in res and head != "fp32" else None),
        "forms": res,
        "roofline": rl,
        "kernels": roof,
        "nufft_rel_l2_vs_nudft": nufft_err,
        "accuracy_reference": acc.get("_reference") if acc else None,
        "oracle_fixture": golden,
        "dropin_host_io": dropin,
        "cpu_baseline": cpu,
        "tuning": {"nufft_tiling": W["nufft_tune"], "gemm_tile_rows": gemm_tile,
                   "gemm_timings": gemm_tune},
        "batching": batching,
        "serving": serving_res,
        "merge": merge_res,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
