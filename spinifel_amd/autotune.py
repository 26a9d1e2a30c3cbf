"""Autotuner for the tile choices that depend on the shapes at hand.

* NUFFT pixel tiling: 16 x 16, 8 x 32, 32 x 8, 8 x 16, 16 x 8, 8 x 8, ...
  detector blocks (256, 128 or 64 pixels).  Larger tiles amortise the
  per-tile staging and barriers over more points; smaller ones keep the
  fine-grid box of a large model (M >= 121) inside LDS.
* Dense top-K GEMM tile: 128 or 256 orientation rows (``mtip_set_gemm_tile``).

Candidates are timed with HIP events on a sample of the real work (the
minimum of a few repetitions after a warm-up) and the winner is cached per
device and shape in a JSON file: ``$SPINIFEL_AMD_TUNE_CACHE``, default
``~/.cache/spinifel_amd/tune.json``.
"""
import json
import os

import numpy as np
import torch

from . import _lib, engine
from .quaternions import random_quaternions

NUFFT_SHAPES = [(16, 16), (8, 32), (32, 8), (8, 16), (16, 8), (8, 8), (4, 16), (16, 4)]


def cache_path():
    return os.environ.get("SPINIFEL_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "spinifel_amd",
                                       "tune.json"))


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    data = _load()
    data[key] = value
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass          # read-only home: tune again next time


def device_key(device):
    p = torch.cuda.get_device_properties(device)
    return f"{p.name}|{getattr(p, 'gcnArchName', '')}|{p.multi_processor_count}"


def time_ms(fn, reps=3, warmup=1):
    """Minimum wall time (ms, HIP events on the current stream) of fn()."""
    for _ in range(warmup):
        fn()
    best = float("inf")
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def nufft_tiling(plan, q, scale, det_shape, n_sample=512, seed=0, use_cache=True,
                 shapes=NUFFT_SHAPES):
    """The fastest Tiling of this detector for this plan (model size, kernel width).

    q: pixel q-vectors (3, N_p) (numpy or device tensor); scale as for slices.
    Returns (engine.Tiling, {shape: ms}).
    """
    det_shape = tuple(int(d) for d in det_shape)
    dev = plan.device
    rows, cols = det_shape[-2:]
    key = f"nufft|{device_key(dev)}|M{plan.M}|w{plan.width}|{det_shape}"
    if use_cache:
        hit = _load().get(key)
        if hit:
            return engine.tiling(det_shape, tuple(hit["tile"])), hit.get("ms", {})
    cands = [s for s in shapes if s[0] * s[1] in engine.TILE_SHAPES and s[0] <= rows
             and s[1] <= cols]
    qd = torch.as_tensor(np.ascontiguousarray(q, np.float64)).to(dev) \
        if isinstance(q, np.ndarray) else q.to(dev, torch.float64)
    quats = torch.as_tensor(random_quaternions(n_sample, np.random.default_rng(seed))).to(dev)
    N_p = qd.shape[1]
    out = torch.empty((n_sample, engine.padded(N_p)), dtype=torch.float32, device=dev)
    times = {}
    for s in cands:
        t = engine.tiling(det_shape, s)
        times[f"{s[0]}x{s[1]}"] = time_ms(lambda: plan.slices(quats, qd, scale, out=out, order=t))
    best = min(times, key=times.get)
    tile = tuple(int(x) for x in best.split("x"))
    _store(key, {"tile": list(tile), "ms": times, "n_sample": n_sample})
    return engine.tiling(det_shape, tile), times


def gemm_tile(precision, S, N_o, D, N_i, N_p, K=8, use_cache=True):
    """Time the 128- and 256-row top-K GEMM tiles on these operands, keep the faster.

    S, D: engine.Operand of slices / images.  Sets the library's tile for
    `precision` and returns (rows, {rows: ms}).
    """
    lib = _lib.load()
    code = _lib.PRECISIONS[precision]
    dev = D.data.device
    key = f"gemm|{device_key(dev)}|{precision}|{N_o}|{N_i}|{N_p}|K{K}"
    if use_cache:
        hit = _load().get(key)
        if hit:
            _lib.check(lib.mtip_set_gemm_tile(code, int(hit["rows"])), "mtip_set_gemm_tile")
            return int(hit["rows"]), hit.get("ms", {})
    one = engine._one(dev)
    ws = torch.empty(max(engine.workspace_bytes(N_o, N_i, K, precision), 1), dtype=torch.uint8,
                     device=dev)
    times = {}
    for rows in (128, 256):
        _lib.check(lib.mtip_set_gemm_tile(code, rows), "mtip_set_gemm_tile")
        times[str(rows)] = time_ms(lambda: engine.topk(precision, S, S.sumsq, N_o, D, N_i, N_p,
                                                       one, K, ws))
    best = int(min(times, key=times.get))
    _lib.check(lib.mtip_set_gemm_tile(code, best), "mtip_set_gemm_tile")
    _store(key, {"rows": best, "ms": times})
    return best, times
