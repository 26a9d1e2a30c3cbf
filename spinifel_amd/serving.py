"""Serving form: resident models, continuous batching of image requests.

A ``Server`` keeps several models resident in HBM, each with its reference
slices computed once (FP32 for rescoring plus the centred residual in the
GEMM format, bf16 by default).  Requests (a batch of images tagged with a
model) are queued; a scheduler thread keeps up to ``n_slots`` batches in
flight on ``compute_streams`` HIP streams.  Whenever a slot is free it takes
the oldest pending request and every other pending request for the same
model (up to ``max_batch_images``), provided the device has enough free HBM
for the batch (``mtip_mem_info`` plus the blocks PyTorch's caching allocator
holds free); requests that arrive while others run are therefore batched as
soon as a slot frees up.  A batch that cannot be admitted while every slot
is idle, even after the allocator's cache is released, fails its requests
with MemoryError after ``admit_timeout_s`` instead of waiting forever.

Image uploads run on one upload stream in admission order, ahead of the
compute streams that wait for them, so the host link stays busy while
admitted batches compute.  Requests handed in as pinned host tensors are
uploaded from where they are, numpy requests through a pinned staging buffer
per slot.  The upload stream never waits on a compute stream, and the
compute streams are few (two by default): HIP maps streams onto a handful of
hardware queues (four unless GPU_MAX_HW_QUEUES says otherwise), and a stream
that shares a queue with a long GEMM has its copies and markers held behind
that kernel.

Each request is scaled by its own std ratio (``mtip_segment_scaling_ratio``)
and every per-image quantity (cross term, top-K keys, FP64 rescore) depends
only on that image; the GEMM tile height is fixed per model (from its
orientation count, not the batch size), so the same kernel scores a request
whatever shares its batch, and a request's result is the same alone and in
any mix, bitwise.
"""
import ctypes
import threading
import time
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, engine


@dataclass
class ServedModel:
    name: str
    ref_quats: np.ndarray
    slices: engine.SliceSet
    operand: engine.Operand
    N_p: int
    det_shape: tuple
    tile: int                     # GEMM tile rows used for every batch of this model


@dataclass(eq=False)
class Request:
    rid: int
    model: str
    images: object                # numpy (N, N_p) float32 or a pinned host tensor
    t_submit: float
    future: Future = field(default_factory=Future)
    n_images: int = 0
    t_start: float = 0.0
    t_done: float = 0.0
    batch_size: int = 0


@dataclass
class Result:
    orientations: np.ndarray
    index: np.ndarray
    dist64: np.ndarray
    latency_s: float
    batch_images: int


def free_hbm(device=0):
    lib = _lib.load()
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.check(lib.mtip_mem_info(int(device), ctypes.byref(f), ctypes.byref(t)), "mtip_mem_info")
    return f.value, t.value


class Server:
    def __init__(self, precision="bf16", K=8, n_slots=4, max_batch_images=4096,
                 mem_reserve_bytes=8 << 30, device=0, eps=1e-6, admit_timeout_s=30.0,
                 stats_window=100000, compute_streams=2):
        engine.require_gpu()
        self.precision = precision
        self.K = K
        self.device = torch.device("cuda", device)
        self.n_slots = n_slots
        self.max_batch_images = max_batch_images
        self.mem_reserve = mem_reserve_bytes
        self.admit_timeout_s = admit_timeout_s
        self._blocked_since = None     # when admission first failed with every slot idle
        self.eps = eps
        self.models = {}
        self._plans = {}
        self._queue = deque()
        self._lock = threading.Condition()
        # batches in flight (slots) share `compute_streams` compute streams
        self._streams = [torch.cuda.Stream(device=self.device)
                         for _ in range(max(1, min(compute_streams, n_slots)))]
        self._upload_stream = torch.cuda.Stream(device=self.device)
        self._busy = [None] * n_slots          # (requests, event, outputs) per slot
        # per slot a pinned host staging buffer, reused while the slot is free:
        # requests are copied into it once and uploaded from it asynchronously
        self._stage = [dict() for _ in range(n_slots + 1)]   # per N_p; the last serves match_now
        self._now_lock = threading.Lock()
        self._next_id = 0
        self.error = None
        self._stop = False
        self._thread = None
        # the last stats_window finished requests (timings only: a request
        # drops its images once its result is set)
        self.completed = deque(maxlen=stats_window)

    # ------------------------------------------------------------- models
    def add_model(self, name, ac, ref_quats, pixel_position_reciprocal, pixel_distance_reciprocal,
                  oversampling=1, tiling=None):
        """Compute and keep the model's reference slices (tiling: an
        engine.Tiling of the detector, default 16 x 16 blocks)."""
        ac = np.asarray(ac, np.float32)
        M = ac.shape[0]
        key = (M, self.eps)
        if key not in self._plans:
            self._plans[key] = engine.NufftPlan(M, self.eps, self.device.index or 0)
        plan = self._plans[key]
        plan.set_model(ac)
        ppr = np.asarray(pixel_position_reciprocal, np.float64)
        E = float(np.asarray(pixel_distance_reciprocal).max())
        order = tiling if tiling is not None else engine.auto_tiling(plan, ppr.shape[1:])
        sl = engine.SliceSet.compute(plan, np.asarray(ref_quats, np.float64), ppr.reshape(3, -1),
                                     np.pi / (E * oversampling), order=order)
        torch.cuda.synchronize(self.device)
        if int(sl.bad.item()) > 0:
            raise AssertionError("NU points outside [-3pi, 3pi)")
        op = sl.operand(self.precision, True)
        # tile rows from the orientation count alone: the 256-row tile for the
        # narrow formats once there are enough orientations to fill it
        tile = 256 if (self.precision != "fp32" and sl.N_o >= 1024) else 128
        self.models[name] = ServedModel(name, np.asarray(ref_quats), sl, op, sl.N_p, ppr.shape[1:],
                                        tile)
        for slot in range(self.n_slots + 1):     # pinned staging allocated up front
            self._stage_buffer(slot, self.max_batch_images, sl.N_p)
        return self.models[name]

    # ------------------------------------------------------------- requests
    def submit(self, model, images):
        """Queue a request; returns a Future of its Result.  images: numpy
        (copied into a pinned staging buffer at launch) or a pinned host
        torch float32 tensor (uploaded from where it is, no host copy; the
        caller must not change it before the future completes)."""
        if model not in self.models:
            raise KeyError(model)
        if torch.is_tensor(images):
            if images.device.type != "cpu" or not images.is_pinned() or \
                    images.dtype != torch.float32:
                raise ValueError("tensor requests must be pinned host float32 tensors")
            images = images.reshape(images.shape[0], -1)
            if not images.is_contiguous():
                # a contiguous copy is not pinned by itself
                images = images.contiguous().pin_memory()
        else:
            images = np.ascontiguousarray(images, np.float32).reshape(len(images), -1)
        if images.shape[1] != self.models[model].N_p:
            raise ValueError(f"images have {images.shape[1]} pixels, model {model} "
                             f"{self.models[model].N_p}")
        with self._lock:
            r = Request(self._next_id, model, images, time.perf_counter(), n_images=len(images))
            self._next_id += 1
            self._queue.append(r)
            self._lock.notify()
        return r.future

    def match_now(self, model, images):
        """Synchronous single-request match (no queue), the 'alone' reference."""
        if not torch.is_tensor(images):
            images = np.ascontiguousarray(images, np.float32).reshape(len(images), -1)
        r = Request(-1, model, images, time.perf_counter(), n_images=len(images))
        with self._now_lock:
            out = self._launch([r], torch.cuda.current_stream(self.device), self.n_slots)
            torch.cuda.synchronize(self.device)
            return self._collect([r], out)[0]

    # ------------------------------------------------------------- engine
    def _batch_bytes(self, n_img, N_p):
        per = N_p * (4 + 4 + 2) + 64 * self.K
        return n_img * per + (64 << 20)

    def _stage_buffer(self, slot, n, n_p):
        buf = self._stage[slot].get(n_p)
        if buf is None or buf.shape[0] < n:
            buf = torch.empty((max(self.max_batch_images, n), n_p), dtype=torch.float32,
                              pin_memory=True)
            self._stage[slot][n_p] = buf
        return buf

    def _upload(self, reqs, slot, n, n_p, stream):
        """The batch's images on the device (n, n_p) float32.

        The copies go on the server's one upload stream, in admission order,
        and `stream` (the slot's compute stream) waits for them: uploads of
        admitted batches run back to back over the host link while earlier
        batches compute.  Pinned tensor requests go up directly; numpy
        requests are first copied into the slot's pinned buffer, which is
        only rewritten once the slot's previous batch has retired (its
        upload completed)."""
        up = self._upload_stream
        buf = None
        a = 0
        with torch.cuda.stream(up):
            # from the upload stream's pool: a block there is freed only when
            # the batch that read it has retired, so the copies can write it
            # at once.  The upload stream never waits on a compute stream
            # (a long kernel elsewhere cannot hold the copies back).
            dev_buf = torch.empty((n, n_p), dtype=torch.float32, device=self.device)
            for r in reqs:
                k = r.n_images
                if torch.is_tensor(r.images):
                    dev_buf[a:a + k].copy_(r.images, non_blocking=True)
                else:
                    if buf is None:
                        buf = self._stage_buffer(slot, n, n_p)
                    # torch's copy runs on its intra-op threads, outside the interpreter lock
                    buf[a:a + k].copy_(torch.from_numpy(r.images))
                    dev_buf[a:a + k].copy_(buf[a:a + k], non_blocking=True)
                a += k
        stream.wait_stream(up)
        return dev_buf

    def _launch(self, reqs, stream, slot=0):
        m = self.models[reqs[0].model]
        sizes = [r.n_images for r in reqs]
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        dev_imgs = self._upload(reqs, slot, int(offs[-1]), m.N_p, stream)
        with torch.cuda.stream(stream):
            # request offsets go up asynchronously from pinned memory: a
            # pageable copy would block this thread until the slot's image
            # upload ahead of it has finished
            offs_dev = torch.from_numpy(offs).pin_memory().to(self.device, non_blocking=True)
            imgs = engine.ImageSet(dev_imgs, self.device, stream)
            alpha = engine.segment_scaling_ratio(imgs.row_sum, imgs.row_sumsq, offs_dev,
                                                 m.slices.row_sum, m.slices.row_sumsq, m.N_p, stream)
            D = imgs.operand(self.precision, alpha, m.slices.center, stream)
            one = engine._one(self.device)
            # one tile height per model and no split edge tiles: an image's
            # keys do not depend on the batch it shares
            idx, val = engine.topk(self.precision, m.operand, m.operand.sumsq, m.slices.N_o, D,
                                   imgs.N_i, imgs.N_p, one, self.K, None, stream, tile=m.tile,
                                   split_edges=False)
            rs = engine.rescore(m.slices.f32, imgs.f32, imgs.N_p, alpha, idx, stream)
            best = rs["best"].to("cpu", non_blocking=True)
            bd = rs["best_d64"].to("cpu", non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        return dict(best=best, bd=bd, event=ev, offs=offs, keep=(offs_dev, imgs, D, idx, rs))

    def _collect(self, reqs, out):
        now = time.perf_counter()
        best = out["best"].numpy().astype(np.int64)
        bd = out["bd"].numpy()
        res = []
        for r, a, b in zip(reqs, out["offs"][:-1], out["offs"][1:]):
            m = self.models[r.model]
            ix = best[a:b]
            res.append(Result(m.ref_quats[ix], ix, bd[a:b], now - r.t_submit, int(out["offs"][-1])))
        return res

    # ------------------------------------------------------------- scheduler
    def start(self):
        self._stop = False
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self._thread.start()

    def stop(self):
        with self._lock:
            self._stop = True
            self._lock.notify()
        if self._thread:
            self._thread.join()

    def _take_batch(self):
        """Oldest pending request plus the other pending ones of its model."""
        if not self._queue:
            return None
        first = self._queue[0]
        m = self.models[first.model]
        batch, n = [], 0
        for r in list(self._queue):
            if r.model != first.model:
                continue
            if batch and n + r.n_images > self.max_batch_images:
                break
            batch.append(r)
            n += r.n_images
        need = self._batch_bytes(n, m.N_p)
        if self._available_bytes() - self.mem_reserve < need:
            if all(b is None for b in self._busy):
                # nothing running will free memory: give the allocator's cache
                # back, and fail the batch if it still does not fit in time
                torch.cuda.empty_cache()
                if self._available_bytes() - self.mem_reserve >= need:
                    self._blocked_since = None
                else:
                    now = time.perf_counter()
                    if self._blocked_since is None:
                        self._blocked_since = now
                    elif now - self._blocked_since > self.admit_timeout_s:
                        self._blocked_since = None
                        err = MemoryError(f"batch of {n} images needs {need} bytes of HBM "
                                          f"beyond the {self.mem_reserve}-byte reserve; "
                                          "none came free")
                        for r in batch:
                            self._queue.remove(r)
                            r.future.set_exception(err)
                    return None
            else:
                return None                  # wait for running batches to free memory
        self._blocked_since = None
        for r in batch:
            self._queue.remove(r)
        return batch

    def _available_bytes(self):
        """Free HBM plus the blocks PyTorch's caching allocator holds unused
        (freed tensors stay reserved by the allocator and out of hipMemGetInfo)."""
        free, _ = free_hbm(self.device.index or 0)
        cached = torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
        return free + max(0, cached)

    def _loop(self):
        try:
            self._loop_body()
        except BaseException as e:          # surface scheduler failures to every waiter
            self.error = e
            with self._lock:
                pending = list(self._queue)
                self._queue.clear()
            for job in self._busy:
                if job is not None:
                    pending.extend(job[0])
            for r in pending:
                if not r.future.done():
                    r.future.set_exception(e)

    def _loop_body(self):
        torch.cuda.set_device(self.device)
        while True:
            progressed = False
            # retire finished slots
            for s in range(self.n_slots):
                job = self._busy[s]
                if job is not None and job[1]["event"].query():
                    reqs, out = job
                    for r, res in zip(reqs, self._collect(reqs, out)):
                        r.t_done = time.perf_counter()
                        r.batch_size = res.batch_images
                        r.images = None
                        self.completed.append(r)
                        r.future.set_result(res)
                    self._busy[s] = None
                    progressed = True
            # admit new batches onto free slots
            with self._lock:
                if self._stop and not self._queue and all(b is None for b in self._busy):
                    return
                for s in range(self.n_slots):
                    if self._busy[s] is None:
                        batch = self._take_batch()
                        if batch is None:
                            break
                        t = time.perf_counter()
                        for r in batch:
                            r.t_start = t
                        stream = self._streams[s % len(self._streams)]
                        self._busy[s] = (batch, self._launch(batch, stream, s))
                        progressed = True
                if not progressed:
                    self._lock.wait(timeout=0.0005)

    # ------------------------------------------------------------- stats
    def reset_stats(self):
        self.completed.clear()

    def stats(self):
        lat = np.array([r.t_done - r.t_submit for r in self.completed])
        n_img = sum(r.n_images for r in self.completed)
        if not len(lat):
            return {}
        t0 = min(r.t_submit for r in self.completed)
        t1 = max(r.t_done for r in self.completed)
        return dict(requests=len(lat), images=n_img, p50_ms=1e3 * float(np.percentile(lat, 50)),
                    p99_ms=1e3 * float(np.percentile(lat, 99)), mean_ms=1e3 * float(lat.mean()),
                    sustained_images_per_s=n_img / (t1 - t0),
                    mean_batch_images=float(np.mean([r.batch_size for r in self.completed])))


def poisson_bursts(rate_hz, duration_s, max_burst, models, rng):
    """Seeded arrival schedule: (time offset, model, burst size) per burst."""
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t > duration_s:
            return out
        out.append((t, models[int(rng.integers(0, len(models)))], int(rng.integers(1, max_burst + 1))))


def run_load(server, schedule, image_source):
    """Replay an arrival schedule in real time; image_source(model, n) -> images."""
    server.start()
    t0 = time.perf_counter()
    futs = []
    for (dt, model, n) in schedule:
        imgs = image_source(model, n)
        while time.perf_counter() - t0 < dt:
            time.sleep(min(0.0005, max(0.0, dt - (time.perf_counter() - t0))))
        futs.append(server.submit(model, imgs))
    for f in futs:
        f.result()
    server.stop()
    return server.stats()
