"""Device-resident orientation-matching pipeline over libmtip.

PyTorch provides device memory and the current HIP stream; every compute
step is a libmtip entry point.  Nothing here falls back to the CPU: a
missing library or GPU raises.

Pipeline of one ``match`` (all on one stream, inputs resident in HBM):

1. ``NufftPlan.set_model``  -- symmetrise / deconvolve / fine grid (once per model)
2. ``NufftPlan.slices``     -- FP64 rotation + ES interpolation -> FP32 slices
                               (+ optional bf16 copy) and per-orientation sums
3. ``ImageSet``             -- images resident as FP32 (+ bf16 / e4m3 copies)
                               with per-row sums
4. ``scaling_ratio``        -- alpha = std(images) / std(slices) on the device
5. ``topk``                 -- MFMA cross term + fused top-K per image
6. ``rescore``              -- FP32 and FP64 full distances of the candidates
"""
import ctypes
import threading
from dataclasses import dataclass
from typing import NamedTuple

import numpy as np
import torch

from . import _lib

_PAD = 64     # row pitch multiple (elements) of every resident operand


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("spinifel_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    _lib.load()


def padded(n, mult=_PAD):
    return ((int(n) + mult - 1) // mult) * mult


class Timers:
    """HIP-event spans on the current stream, keyed by name.

    ``with timers.span("nufft"): ...`` records an event pair around the
    enclosed launches; ``ms(name)`` (after a synchronize) lists durations.
    """

    def __init__(self):
        self.spans = {}

    def span(self, name):
        timers = self

        class _Span:
            def __enter__(self_):
                self_.a = torch.cuda.Event(enable_timing=True)
                self_.b = torch.cuda.Event(enable_timing=True)
                self_.a.record()
                return self_

            def __exit__(self_, *exc):
                self_.b.record()
                timers.spans.setdefault(name, []).append((self_.a, self_.b))
                return False

        return _Span()

    def ms(self, name):
        return [a.elapsed_time(b) for a, b in self.spans.get(name, [])]

    def clear(self):
        self.spans = {}


class _NoSpan:
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


def _span(timers, name):
    return timers.span(name) if timers is not None else _NoSpan()


class Tiling(NamedTuple):
    """Pixel visiting order for the slicing kernel: runs of `points` entries,
    each one tile of neighbouring detector pixels, -1 padded."""
    order: np.ndarray
    points: int
    shape: tuple


TILE_SHAPES = {256: (16, 16), 128: (8, 16), 64: (8, 8)}


def tile_order(det_shape, tile=(16, 16), group=(2, 2)):
    """Pixel visiting order in runs of ty * tx: tile blocks of each panel, -1 padded.

    det_shape (..., rows, cols) with leading dims as panels; returns an int32
    numpy array whose length is a multiple of ty * tx, or None for a 1-D detector.
    Inside a full tile the pixels go in gy x gx groups (row-major groups,
    row-major inside a group; ``group=None`` keeps the tile row-major).  The
    slicing kernel gives four consecutive slots to one quarter-wave (16 lanes)
    of an LDS read, so a 2 x 2 group puts four detector neighbours, whose
    interpolation windows overlap most, on the same read: more lanes share
    dwords and fewer collide on banks.  At M = 81, 10k orientations x 16k
    pixels in 16 x 16 tiles the slicing takes 8.97 ms this way against 9.47 ms
    with row-major tiles (2 x 4 groups: 9.38 ms; 4 x 2: 8.98 ms).
    """
    det_shape = tuple(int(d) for d in det_shape)
    if len(det_shape) < 2:
        return None
    rows, cols = det_shape[-2:]
    panels = int(np.prod(det_shape[:-2])) if len(det_shape) > 2 else 1
    ty, tx = tile
    idx = np.arange(panels * rows * cols, dtype=np.int64).reshape(panels, rows, cols)
    runs = []
    for p in range(panels):
        for r in range(0, rows, ty):
            for c in range(0, cols, tx):
                blk = idx[p, r:r + ty, c:c + tx]
                if group is not None and blk.shape == (ty, tx) and ty % group[0] == 0 \
                        and tx % group[1] == 0:
                    gy, gx = group
                    blk = blk.reshape(ty // gy, gy, tx // gx, gx).transpose(0, 2, 1, 3)
                blk = blk.ravel()
                run = np.full(ty * tx, -1, dtype=np.int64)
                run[:blk.size] = blk
                runs.append(run)
    return np.concatenate(runs).astype(np.int32)


def tiling(det_shape, tile=(16, 16)):
    """A Tiling of the detector in ty x tx blocks (ty * tx in 64, 128, 256)."""
    ty, tx = tile
    if ty * tx not in TILE_SHAPES:
        raise ValueError(f"tile {tile}: ty * tx must be 64, 128 or 256")
    o = tile_order(det_shape, tile)
    return None if o is None else Tiling(o, ty * tx, (int(ty), int(tx)))


def _orient_order(quats):
    """Morton order of the orientations of a device quaternion tensor, kept on
    the tensor object with its version counter (an address-based cache would
    hand a reused allocation the order of other quaternions)."""
    hit = getattr(quats, "_mtip_orient_order", None)
    if hit is not None and hit[0] == quats._version:
        return hit[1]
    from .quaternions import morton_order
    perm = torch.as_tensor(morton_order(quats.cpu().numpy()).astype(np.int32)).to(quats.device)
    try:
        quats._mtip_orient_order = (quats._version, perm)
    except AttributeError:
        pass
    return perm


def _orient_rank(quats):
    """Position of each orientation in ``_orient_order(quats)`` (int32),
    cached on the tensor like the order."""
    hit = getattr(quats, "_mtip_orient_rank", None)
    if hit is not None and hit[0] == quats._version:
        return hit[1]
    perm = _orient_order(quats).long()
    rank = torch.empty(perm.numel(), dtype=torch.int32, device=perm.device)
    rank[perm] = torch.arange(perm.numel(), dtype=torch.int32, device=perm.device)
    try:
        quats._mtip_orient_rank = (quats._version, rank)
    except AttributeError:
        pass
    return rank


def auto_tiling(plan, det_shape):
    """Tiling by rule, without timing: 16 x 16 blocks while a block's fine-grid
    box fits LDS (fine grids up to ~200 points per edge at 0.9 Nyquist), 8 x 8
    blocks beyond.  The autotuner (spinifel_amd.autotune) picks by timing."""
    if len(tuple(det_shape)) < 2:
        return None
    return tiling(det_shape, (16, 16) if plan.n_fine <= 200 else (8, 8))


def auto_spread_tiling(plan, det_shape):
    """Tiling for the type-1 spreading, by rule: 8 x 16 blocks while their
    LDS box fits (fine grids up to ~200 points per edge), 8 x 8 beyond
    (measured at 16.4M points: 8.7 ms for 8 x 16 vs 11.0 for 16 x 16 at
    n = 162; 12.0 ms for 8 x 8 vs 59 for 8 x 16 at n = 324)."""
    if len(tuple(det_shape)) < 2:
        return None
    return tiling(det_shape, (8, 16) if plan.n_fine <= 200 else (8, 8))


class NufftPlan:
    """Type-2 NUFFT plan of a cubic M^3 model (ES kernel, sigma = 2)."""

    def __init__(self, M, eps=1e-6, device=0):
        require_gpu()
        self.lib = _lib.load()
        self.M = int(M)
        self.eps = float(eps)
        self.device = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
        h = ctypes.c_void_p()
        _lib.check(self.lib.mtip_plan_create(self.M, self.eps, self.device.index or 0,
                                             ctypes.byref(h)), "mtip_plan_create")
        self._h = h
        n, w, np_ = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        beta = ctypes.c_double()
        _lib.check(self.lib.mtip_plan_info(h, ctypes.byref(n), ctypes.byref(w),
                                           ctypes.byref(beta), ctypes.byref(np_)), "mtip_plan_info")
        self.n_fine, self.width, self.beta, self.padded_edge = n.value, w.value, beta.value, np_.value
        self._model = None

    def close(self):
        if getattr(self, "_h", None):
            self.lib.mtip_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_model(self, ac, stream=None):
        """ac: (M, M, M) real array or tensor; copied to the device as float32."""
        if isinstance(ac, np.ndarray):
            if np.iscomplexobj(ac):
                if not np.all(np.isreal(ac)):
                    raise AssertionError("model must be real (use_recip_sym)")
                ac = ac.real
            t = torch.from_numpy(np.ascontiguousarray(ac, dtype=np.float32)).to(self.device)
        else:
            t = ac.to(device=self.device, dtype=torch.float32).contiguous()
        if tuple(t.shape) != (self.M,) * 3:
            raise ValueError(f"model shape {tuple(t.shape)} != {(self.M,) * 3}")
        self._model = t
        _lib.check(self.lib.mtip_plan_set_model(self._h, _ptr(t), _stream(stream)),
                   "mtip_plan_set_model")

    def execute(self, x, y, z, stream=None, check_domain=True):
        """Re of the type-2 transform at points (x, y, z) (radians) -> float32 tensor."""
        xs = [torch.as_tensor(v, dtype=torch.float64).reshape(-1).to(self.device).contiguous()
              for v in (x, y, z)]
        N = xs[0].numel()
        if not (xs[1].numel() == N and xs[2].numel() == N):
            raise ValueError("x, y, z differ in length")
        out = torch.empty(N, dtype=torch.float32, device=self.device)
        bad = torch.zeros(1, dtype=torch.int32, device=self.device)
        _lib.check(self.lib.mtip_plan_execute(self._h, _ptr(xs[0]), _ptr(xs[1]), _ptr(xs[2]), N,
                                              _ptr(out), _ptr(bad), _stream(stream)),
                   "mtip_plan_execute")
        if check_domain and int(bad.item()) > 0:
            raise AssertionError(f"{int(bad.item())} NU points outside [-3pi, 3pi)")
        return out

    def slices(self, quats, q, scale, out=None, ld=None, want_bf16=False, bad=None, stream=None,
               timers=None, order=None, orient_order="auto"):
        """Slices at N_o orientations for pixel q-vectors q (3, N_p).

        quats: (N_o, 4) float64 device tensor; q: (3, N_p) float64 device tensor.
        Returns dict(s32 (N_o, ld) float32, bf16 (N_o, ld) int16 bits or None,
        row_sum, row_sumsq (N_o,) float64, bad (1,) int32).  Pixel columns
        beyond N_p are zero.  order: a Tiling (or pixel order in runs of 256);
        orient_order: processing order of the orientations ('auto': Morton
        order of their rotation vectors, cached per quaternion buffer; None:
        index order; or a device int32 permutation).
        """
        quats = quats.to(self.device, torch.float64).contiguous()
        q = q.to(self.device, torch.float64).contiguous()
        N_o = quats.shape[0]
        N_p = q.shape[1]
        tile_points = 256
        if isinstance(order, Tiling):
            order, tile_points = order.order, order.points
        if order is not None:
            if isinstance(order, np.ndarray):
                order = torch.as_tensor(order)
            order = order.to(self.device, torch.int32).contiguous()
        if ld is None:
            ld = padded(N_p)
        if out is None:
            # the kernel writes every pixel column; only pitch padding needs zeros
            out = torch.empty((N_o, ld), dtype=torch.float32, device=self.device)
            if ld > N_p:
                out[:, N_p:].zero_()
        bf = None
        if want_bf16:
            bf = torch.empty((N_o, ld), dtype=torch.int16, device=self.device)
            if ld > N_p:
                bf[:, N_p:].zero_()
        rs = torch.empty(N_o, dtype=torch.float64, device=self.device)
        rq = torch.empty(N_o, dtype=torch.float64, device=self.device)
        if bad is None:
            bad = torch.zeros(1, dtype=torch.int32, device=self.device)
        if isinstance(orient_order, str):
            orient_order = _orient_order(quats) if N_o >= 64 else None
        with _span(timers, "nufft"):
            rc = self.lib.mtip_plan_slices(self._h, _ptr(quats), N_o, _ptr(q), N_p, _ptr(order),
                                           order.numel() if order is not None else 0, tile_points,
                                           _ptr(orient_order), float(scale),
                                           _ptr(out), ld, _ptr(bf), _ptr(rs), _ptr(rq), _ptr(bad),
                                           _stream(stream))
        _lib.check(rc, "mtip_plan_slices")
        return dict(s32=out, bf16=bf, row_sum=rs, row_sumsq=rq, bad=bad, N_p=N_p, ld=ld)

    # ---------------------------------------------------------- type 1 (adjoint)

    def _support_arg(self, support):
        if support is None:
            return None
        s = torch.as_tensor(support)
        if s.dim() == 0 or s.numel() == 1:
            v = float(s.reshape(-1)[0])
            return None if v == 1.0 else torch.full((self.M,) * 3, v, dtype=torch.float64,
                                                    device=self.device)
        s = s.to(self.device, torch.float64).contiguous()
        if tuple(s.shape) != (self.M,) * 3 and s.numel() != self.M ** 3:
            raise ValueError(f"support shape {tuple(s.shape)} != {(self.M,) * 3}")
        return s

    def _adjoint_finish(self, support, out_scale, out, bad, check_domain, stream):
        sup = self._support_arg(support)
        if out is None:
            out = torch.empty((self.M,) * 3, dtype=torch.float64, device=self.device)
        scale = 1.0 / self.M ** 3 if out_scale is None else float(out_scale)
        _lib.check(self.lib.mtip_plan_spread_finish(self._h, _ptr(sup), scale, _ptr(out),
                                                    _stream(stream)), "mtip_plan_spread_finish")
        if check_domain and int(bad.item()) > 0:
            raise AssertionError(f"{int(bad.item())} NU points outside [-3pi, 3pi)")
        return out

    def adjoint_slices(self, quats, q, scale, values=None, support=None, out_scale=None,
                       order=None, out=None, stream=None, check_domain=True):
        """Re of the type-1 transform (isign +1) of ``values`` at the rotated
        pixel q-points -- the adjoint of ``slices``:

            out = out_scale * support * Re sum_j c_j exp(+i k . x_j),  (M, M, M) float64

        with x_j = scale * inv(R(quats[n])) q_p; values (N_o, >= N_p) float32
        (row n at orientation n, indexed by pixel) or None for ones;
        out_scale defaults to 1 / M^3 (spinifel/autocorrelation.py:128).
        order: a Tiling (its tiles share one LDS box; ``auto_tiling`` picks
        16 x 16 or 8 x 8 by grid size) or None (pixel index order, 256 at a time).
        """
        quats = quats.to(self.device, torch.float64).contiguous()
        q = q.to(self.device, torch.float64).contiguous()
        N_o, N_p = quats.shape[0], q.shape[1]
        tile_points = 256
        if isinstance(order, Tiling):
            order, tile_points = order.order, order.points
        if order is not None:
            order = torch.as_tensor(order).to(self.device, torch.int32).contiguous()
        ld = 0
        if values is not None:
            values = torch.as_tensor(values).to(self.device, torch.float32)
            if values.dim() == 1:
                values = values.reshape(N_o, -1)
            if values.stride(-1) != 1:
                values = values.contiguous()
            if values.shape[0] != N_o or values.shape[1] < N_p:
                raise ValueError(f"values shape {tuple(values.shape)} for {N_o} x {N_p} points")
            ld = values.stride(0)
        bad = torch.zeros(1, dtype=torch.int32, device=self.device)
        st = _stream(stream)
        _lib.check(self.lib.mtip_plan_spread_begin(self._h, st), "mtip_plan_spread_begin")
        _lib.check(self.lib.mtip_plan_spread_slices(
            self._h, _ptr(quats), N_o, _ptr(q), N_p, _ptr(order),
            order.numel() if order is not None else 0, tile_points, float(scale), _ptr(values), ld,
            _ptr(bad), st), "mtip_plan_spread_slices")
        return self._adjoint_finish(support, out_scale, out, bad, check_domain, stream)

    def adjoint_points(self, x, y, z, values=None, support=None, out_scale=None, out=None,
                       stream=None, check_domain=True, spatial="auto"):
        """Re of the type-1 transform (isign +1) of values at points (x, y, z)
        (radians) -> (M, M, M) float64, times support and out_scale (1 / M^3).

        spatial: put the points in Morton order of 4-cell cubes of the fine
        grid before spreading ('auto': from 2^20 points on), so that the
        kernel's runs of 128 points fill compact LDS boxes whatever order the
        caller's points come in (orientation-major detector rows make long,
        thin boxes).  The result is the same sum, added in another order."""
        xs = [torch.as_tensor(v, dtype=torch.float64).reshape(-1).to(self.device).contiguous()
              for v in (x, y, z)]
        N = xs[0].numel()
        if not (xs[1].numel() == N and xs[2].numel() == N):
            raise ValueError("x, y, z differ in length")
        if values is not None:
            values = torch.as_tensor(values).reshape(-1).to(self.device, torch.float32).contiguous()
            if values.numel() != N:
                raise ValueError(f"{values.numel()} values for {N} points")
        if spatial == "auto":
            spatial = N >= (1 << 20)
        if spatial:
            perm = spatial_order(xs[0], xs[1], xs[2], 4 * 2 * np.pi / self.n_fine)
            xs = [v[perm] for v in xs]
            if values is not None:
                values = values[perm]
        bad = torch.zeros(1, dtype=torch.int32, device=self.device)
        st = _stream(stream)
        _lib.check(self.lib.mtip_plan_spread_begin(self._h, st), "mtip_plan_spread_begin")
        _lib.check(self.lib.mtip_plan_spread_points(self._h, _ptr(xs[0]), _ptr(xs[1]), _ptr(xs[2]),
                                                    N, _ptr(values), _ptr(bad), st),
                   "mtip_plan_spread_points")
        return self._adjoint_finish(support, out_scale, out, bad, check_domain, stream)


def rotated_points(quats, q, scale, stream=None):
    """x, y, z (radians, FP64 device vectors of N_o * N_p, orientation-major)
    of the pixel q-vectors (3, N_p) rotated by inv(R(quat)) and scaled."""
    lib = _lib.load()
    N_o, N_p = quats.shape[0], q.shape[1]
    outs = [torch.empty(N_o * N_p, dtype=torch.float64, device=q.device) for _ in range(3)]
    st = _stream(stream)
    for s in range(0, N_o, 65535):
        e = min(N_o, s + 65535)
        _lib.check(lib.mtip_rotate_positions(
            ctypes.c_void_p(quats[s:].data_ptr()), e - s, _ptr(q), N_p, float(scale),
            ctypes.c_void_p(outs[0][s * N_p:].data_ptr()),
            ctypes.c_void_p(outs[1][s * N_p:].data_ptr()),
            ctypes.c_void_p(outs[2][s * N_p:].data_ptr()), st), "mtip_rotate_positions")
    return outs


def _part1by2(v):
    # the low 10 bits of v spread to every third bit (bit interleaving)
    v = v & 0x3FF
    v = (v | (v << 16)) & 0x030000FF
    v = (v | (v << 8)) & 0x0300F00F
    v = (v | (v << 4)) & 0x030C30C3
    v = (v | (v << 2)) & 0x09249249
    return v


def spatial_order(x, y, z, cell):
    """Permutation (int64) visiting points (radians, |x| < 3 pi) in Morton
    order of the cubes of edge `cell` radians they fall in: runs of
    consecutive points then share small boxes of the fine grid, so the
    spreading kernel's 128-point tiles accumulate compact boxes in LDS."""
    idx = [torch.clamp(((c + 3 * np.pi) / cell).floor().to(torch.int32), 0, 1023)
           for c in (x, y, z)]
    key = _part1by2(idx[0]) | (_part1by2(idx[1]) << 1) | (_part1by2(idx[2]) << 2)
    # stable: points of one cube keep the caller's order, so each 128-point
    # run holds the same points in the same order on every call and fills its
    # LDS box identically; the runs' boxes still reach the FP64 grid through
    # global atomics, whose order varies, so the grid is reproducible to
    # rounding, not bitwise
    return torch.argsort(key, stable=True)


def row_stats(x, n_cols=None, stream=None):
    """FP64 per-row sum and sum of squares of a float32 (rows, ld) tensor."""
    lib = _lib.load()
    rows, ld = x.shape
    cols = ld if n_cols is None else n_cols
    rs = torch.empty(rows, dtype=torch.float64, device=x.device)
    rq = torch.empty(rows, dtype=torch.float64, device=x.device)
    _lib.check(lib.mtip_row_stats(_ptr(x), rows, cols, ld, _ptr(rs), _ptr(rq), _stream(stream)),
               "mtip_row_stats")
    return rs, rq


def quantize(x, fmt, n_cols=None, stream=None):
    """Cast (rows, ld) float32 rows to bf16 bits (int16) or e4m3 (uint8 + scales)."""
    lib = _lib.load()
    rows, ld = x.shape
    cols = ld if n_cols is None else n_cols
    if fmt == "bf16":
        out = torch.empty((rows, ld), dtype=torch.int16, device=x.device)
        sc = None
        code = _lib.MTIP_BF16
    elif fmt == "fp8":
        out = torch.empty((rows, ld), dtype=torch.uint8, device=x.device)
        sc = torch.empty(rows, dtype=torch.float32, device=x.device)
        code = _lib.MTIP_FP8
    else:
        raise ValueError(fmt)
    _lib.check(lib.mtip_quantize_rows(_ptr(x), rows, cols, ld, code, _ptr(out), ld, _ptr(sc),
                                      _stream(stream)), "mtip_quantize_rows")
    return out, sc


def scaling_ratio(img_sum, img_sumsq, sl_sum, sl_sumsq, n_pix, stream=None):
    lib = _lib.load()
    alpha = torch.empty(1, dtype=torch.float64, device=img_sum.device)
    _lib.check(lib.mtip_scaling_ratio(_ptr(img_sum), _ptr(img_sumsq), img_sum.numel(),
                                      _ptr(sl_sum), _ptr(sl_sumsq), sl_sum.numel(), int(n_pix),
                                      _ptr(alpha), _stream(stream)), "mtip_scaling_ratio")
    return alpha


def segment_scaling_ratio(img_sum, img_sumsq, offsets, sl_sum, sl_sumsq, n_pix, stream=None):
    """alpha per image: each segment (request) scaled by its own image std."""
    lib = _lib.load()
    dev = img_sum.device
    if torch.is_tensor(offsets) and offsets.device == dev and offsets.dtype == torch.int64:
        off = offsets.contiguous()       # already resident (no copy, capturable)
    else:
        off = torch.as_tensor(np.asarray(offsets, dtype=np.int64)).to(dev)
    alpha = torch.empty(img_sum.numel(), dtype=torch.float64, device=dev)
    _lib.check(lib.mtip_segment_scaling_ratio(_ptr(img_sum), _ptr(img_sumsq), _ptr(off),
                                              off.numel() - 1, _ptr(sl_sum), _ptr(sl_sumsq),
                                              sl_sum.numel(), int(n_pix), _ptr(alpha),
                                              _stream(stream)), "mtip_segment_scaling_ratio")
    return alpha


@dataclass
class Operand:
    """A row matrix resident in one GEMM operand format: data (rows, ld),
    per-row dequantisation scales (e4m3 only) and FP64 squared row norms of
    the values as cast."""
    data: torch.Tensor
    scale: torch.Tensor = None
    sumsq: torch.Tensor = None


_FMT = {"fp32": (_lib.MTIP_FP32, torch.float32), "bf16": (_lib.MTIP_BF16, torch.int16),
        "fp8": (_lib.MTIP_FP8, torch.uint8)}


def residual(x, precision, n_cols=None, mult=None, invert=False, center=None, stream=None):
    """Operand of rows ``cast(m * x - center)`` (m = mult or 1 / mult, a device scalar)."""
    lib = _lib.load()
    rows, ld = x.shape
    cols = ld if n_cols is None else n_cols
    code, dt = _FMT[precision]
    out = torch.empty((rows, ld), dtype=dt, device=x.device)
    sc = torch.empty(rows, dtype=torch.float32, device=x.device) if precision == "fp8" else None
    ss = torch.empty(rows, dtype=torch.float64, device=x.device)
    flags = (1 if invert else 0) | (2 if (mult is not None and mult.numel() > 1) else 0)
    _lib.check(lib.mtip_residual_rows(_ptr(x), rows, cols, ld, _ptr(mult), flags,
                                      _ptr(center), code, _ptr(out), ld, _ptr(sc), _ptr(ss),
                                      _stream(stream)), "mtip_residual_rows")
    return Operand(out, sc, ss)


def col_mean(x, n_cols=None, stream=None, row_step=1):
    """Column means of (rows, ld) float32 rows -> (ld,) float32 (zeros past n_cols),
    over every row_step-th row (rows 0, row_step, 2 row_step, ...)."""
    lib = _lib.load()
    rows, ld = x.shape
    cols = ld if n_cols is None else n_cols
    row_step = max(1, int(row_step))
    out = torch.zeros(ld, dtype=torch.float32, device=x.device)
    nb = int(lib.mtip_col_mean_workspace_size(cols))
    work = torch.empty(max(nb, 8), dtype=torch.uint8, device=x.device)
    _lib.check(lib.mtip_col_mean(_ptr(x), (rows + row_step - 1) // row_step, cols, ld * row_step,
                                 _ptr(work), work.numel(), _ptr(out), _stream(stream)),
               "mtip_col_mean")
    return out


# rows the centre of a slice set is averaged over (at least): any common vector
# leaves the distances unchanged, and the mean of about a thousand orientations
# already removes the orientation-independent part of the pattern
_CENTER_ROWS = 1024


# pinned host staging and a copy stream per device for upload_rows (the lock
# serialises callers on different threads: the buffer is shared)
_STAGING = {}
_UPLOAD_STREAMS = {}
_STAGING_LOCK = threading.Lock()


def upload_rows(rows, device, chunk_rows=1024):
    """Host rows (N, P) -> device float32 tensor (N, P), uploaded on a copy
    stream of its own while the current stream keeps running.

    The rows go through one pinned staging buffer kept per device (grown to
    the largest upload and reused; ``release_staging`` frees it), chunk by
    chunk: the host copy of a chunk overlaps the DMA of the one before, and
    no chunk waits on the current stream, so an upload issued after a kernel
    runs beside it.  The current stream waits for the whole upload before
    anything it issues later.  The device tensor comes from the copy
    stream's pool and is marked in use by the current stream.
    """
    dev = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
    arr = np.ascontiguousarray(rows, dtype=np.float32)
    arr = arr.reshape(arr.shape[0], -1)
    with _STAGING_LOCK:
        return _upload_rows_locked(arr, dev, chunk_rows)


def _upload_rows_locked(arr, dev, chunk_rows):
    n, p = arr.shape
    cur = torch.cuda.current_stream(dev)
    up = _UPLOAD_STREAMS.get(dev)
    if up is None:
        up = _UPLOAD_STREAMS[dev] = torch.cuda.Stream(device=dev)
    st = _STAGING.get(dev)
    if st is None or st["buf"].numel() < arr.size:
        if st is not None:
            st["done"].synchronize()
        st = _STAGING[dev] = dict(buf=torch.empty(max(arr.size, 1), dtype=torch.float32,
                                                  pin_memory=True),
                                  done=torch.cuda.Event())
    else:
        st["done"].synchronize()        # the previous upload from the buffer has finished
    with torch.cuda.stream(up):
        out = torch.empty((n, p), dtype=torch.float32, device=dev)
    out.record_stream(cur)
    host = st["buf"][:arr.size].view(n, p)
    src = torch.from_numpy(arr)
    step = max(1, int(chunk_rows))
    with torch.cuda.stream(up):
        for a in range(0, n, step):
            b = min(n, a + step)
            host[a:b].copy_(src[a:b])   # torch's threaded copy, outside the interpreter lock
            out[a:b].copy_(host[a:b], non_blocking=True)
        st["done"].record(up)
    cur.wait_event(st["done"])
    return out


def release_staging():
    """Free upload_rows' pinned staging buffers (after their uploads finish)."""
    with _STAGING_LOCK:
        for st in _STAGING.values():
            st["done"].synchronize()
        _STAGING.clear()


class ImageSet:
    """Images resident in HBM: FP32 rows padded to a multiple of 16 pixels,
    FP64 per-row sums, and lazily built bf16 / e4m3 copies."""

    def __init__(self, images, device=0, stream=None):
        require_gpu()
        dev = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
        if isinstance(images, np.ndarray):
            arr = np.ascontiguousarray(images, dtype=np.float32).reshape(images.shape[0], -1)
            t = torch.from_numpy(arr).to(dev)
        else:
            t = images.to(dev, torch.float32).reshape(images.shape[0], -1)
        self.N_i, self.N_p = t.shape
        self.ld = padded(self.N_p)
        if self.ld != self.N_p:
            buf = torch.zeros((self.N_i, self.ld), dtype=torch.float32, device=dev)
            buf[:, :self.N_p] = t
            t = buf
        self.f32 = t.contiguous()
        self.device = dev
        self.row_sum, self.row_sumsq = row_stats(self.f32, self.N_p, stream)
        self._ops = {}

    def operand(self, precision, alpha, center=None, stream=None):
        """Images divided by alpha minus the slice-side centre, cast to `precision`."""
        return residual(self.f32, precision, self.N_p, mult=alpha, invert=True, center=center,
                        stream=stream)


_SPLIT_EDGES = 1      # MTIP_MATCH_SPLIT_EDGES


def workspace_bytes(N_o, N_i, K, precision):
    return int(_lib.load().mtip_match_workspace_size(N_o, N_i, K, _lib.PRECISIONS[precision]))


def topk(precision, S, s_sumsq, N_o, D, N_i, N_p, alpha, K, workspace=None, stream=None,
         timers=None, tile=None, split_edges=True):
    """Top-K orientations per image by v = alpha^2 ||s||^2 - 2 alpha s.d.

    S, D: Operand; returns (idx int32 (N_i, K), val float32 (N_i, K)).
    tile: GEMM tile rows (128 or 256); None takes mtip_gemm_tile's choice for
    these sizes.  split_edges: let the 256-row kernel run thin edge tiles as
    split-K parts (MTIP_MATCH_SPLIT_EDGES); False keeps one summation order
    for every tile, so that an image's keys do not depend on where the batch
    puts it.
    """
    lib = _lib.load()
    dev = D.data.device
    need = workspace_bytes(N_o, N_i, K, precision)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(max(need, 1), dtype=torch.uint8, device=dev)
    idx = torch.empty((N_i, K), dtype=torch.int32, device=dev)
    val = torch.empty((N_i, K), dtype=torch.float32, device=dev)
    st = _stream(stream)
    code = _lib.PRECISIONS[precision]
    if tile is None:
        # one tile height for both halves, so that a concurrent
        # mtip_set_gemm_tile (autotuner, another thread) cannot change it
        # between them
        tile = lib.mtip_gemm_tile(code, N_o, N_i)
        _lib.check(min(tile, 0), "mtip_gemm_tile")
    with _span(timers, "gemm_" + precision):
        rc = lib.mtip_match_partial_ex(code, _ptr(S.data), S.data.shape[1],
                                       _ptr(S.scale), _ptr(s_sumsq), N_o, _ptr(D.data),
                                       D.data.shape[1], _ptr(D.scale), N_i, N_p, _ptr(alpha), K,
                                       tile, _SPLIT_EDGES if split_edges else 0,
                                       _ptr(workspace), workspace.numel(), st)
    _lib.check(rc, "mtip_match_partial_ex")
    with _span(timers, "merge"):
        rc = lib.mtip_topk_merge(code, _ptr(workspace), N_o, N_i, K, tile,
                                 _ptr(idx), _ptr(val), st)
    _lib.check(rc, "mtip_topk_merge")
    return idx, val


def rescore(S32, D32, N_p, alpha, cand, stream=None, timers=None, image_order=None):
    """FP32 and FP64 distances of the candidates; FP64 argmin per image.
    image_order: optional device int32 processing order of the images."""
    lib = _lib.load()
    N_i, K = cand.shape
    dev = cand.device
    best = torch.empty(N_i, dtype=torch.int32, device=dev)
    bd = torch.empty(N_i, dtype=torch.float64, device=dev)
    d32 = torch.empty((N_i, K), dtype=torch.float32, device=dev)
    d64 = torch.empty((N_i, K), dtype=torch.float64, device=dev)
    with _span(timers, "rescore"):
        rc = lib.mtip_rescore(_ptr(S32), S32.shape[1], S32.shape[0], _ptr(D32), D32.shape[1], N_i,
                              N_p, _ptr(alpha), int(alpha.numel() > 1), _ptr(cand.contiguous()), K,
                              _ptr(best), _ptr(bd),
                              _ptr(d32), _ptr(d64), _ptr(image_order), _stream(stream))
    _lib.check(rc, "mtip_rescore")
    return dict(best=best, best_d64=bd, d32=d32, d64=d64)


def euclidean_distance(ref, data, stream=None):
    """Squared distances (numData, numRef) float32 (cudaEuclideanDistance)."""
    lib = _lib.load()
    ref = ref.contiguous()
    data = data.contiguous()
    numRef, P = ref.shape
    numData = data.shape[0]
    out = torch.empty((numData, numRef), dtype=torch.float32, device=data.device)
    _lib.check(lib.mtip_euclidean_distance(_ptr(ref), _ptr(data), numRef, numData, P, _ptr(out),
                                           _stream(stream)), "mtip_euclidean_distance")
    return out


def heap_sort(euDist, totalNN, stream=None):
    """Indices of the totalNN smallest per row (cudaHeapSort)."""
    lib = _lib.load()
    euDist = euDist.contiguous()
    nd, nr = euDist.shape
    idx = torch.empty((nd, totalNN), dtype=torch.int32, device=euDist.device)
    _lib.check(lib.mtip_heap_sort(_ptr(euDist), nd, nr, int(totalNN), _ptr(idx), _stream(stream)),
               "mtip_heap_sort")
    return idx


class SliceSet:
    """Slices of one model at one orientation set, resident in HBM as FP32
    rows (N_o, ld) with FP64 per-row sums, plus lazily built bf16 / e4m3
    operand copies."""

    def __init__(self, f32, N_p, row_sum, row_sumsq, bf16=None, quats=None, bad=None):
        self.f32 = f32
        self.N_o = f32.shape[0]
        self.N_p = int(N_p)
        self.row_sum, self.row_sumsq = row_sum, row_sumsq
        self.quats = quats
        self.bad = bad
        self.bf16_plain = bf16
        self._ops = {}
        self._center = None

    @classmethod
    def compute(cls, plan, quats, q, scale, precision="fp32", stream=None, timers=None, order=None):
        """NUFFT slices of the plan's model at quaternions `quats` (N_o, 4)."""
        dev = plan.device
        quats = torch.as_tensor(np.ascontiguousarray(quats, np.float64)).to(dev) \
            if isinstance(quats, np.ndarray) else quats.to(dev, torch.float64)
        q = torch.as_tensor(np.ascontiguousarray(q, np.float64)).to(dev) \
            if isinstance(q, np.ndarray) else q.to(dev, torch.float64)
        r = plan.slices(quats, q, scale, stream=stream, timers=timers, order=order)
        return cls(r["s32"], r["N_p"], r["row_sum"], r["row_sumsq"], r["bf16"], quats, r["bad"])

    @classmethod
    def from_rows(cls, rows, N_p=None, stream=None):
        """Wrap given FP32 rows (N_o, ld) device tensor (ld a multiple of 16)."""
        N_p = rows.shape[1] if N_p is None else N_p
        rs, rq = row_stats(rows, N_p, stream)
        return cls(rows, N_p, rs, rq)

    @property
    def center(self):
        """Per-pixel mean slice (float32, ld), the centre both operands subtract:
        the mean of every k-th orientation's slice, k = N_o // 1024 (all of
        them below 2048 orientations).  Distances are invariant to the choice
        of centre; it only has to remove the orientation-independent part of
        the pattern before the cast to bf16 / e4m3."""
        if self._center is None:
            self._center = col_mean(self.f32, self.N_p,
                                    row_step=max(1, self.N_o // _CENTER_ROWS))
        return self._center

    @property
    def rank(self):
        """Position of each orientation in the Morton order of the slicing
        (int32, N_o), or None when the orientations are not known."""
        if self.quats is None or self.N_o < 64:
            return None
        return _orient_rank(self.quats)

    def operand(self, precision, center=True, stream=None):
        """Slices minus the centre (or plain), cast to `precision`, cached."""
        key = (precision, bool(center))
        if key not in self._ops:
            c = self.center if center else None
            self._ops[key] = residual(self.f32, precision, self.N_p, center=c, stream=stream)
        return self._ops[key]


_ONES = {}


def _one(dev):
    if dev not in _ONES:
        _ONES[dev] = torch.ones(1, dtype=torch.float64, device=dev)
    return _ONES[dev]


def match_resident(slices, images, precision="fp32", K=None, rescore_fp64=None, alpha=None,
                   center=True, workspace=None, stream=None, timers=None):
    """Match resident images against resident slices.

    Ranking uses the residuals s' = s - c and d' = d / alpha - c with c the
    mean slice (``center=True``): ||s' - d'|| = ||alpha s - d|| / alpha, so
    the order is the order of the true distances, while the cast to bf16 /
    e4m3 only sees the orientation-dependent part.  precision 'fp32'
    defaults to K = 1 without rescoring (the plain form); 'bf16' / 'fp8'
    default to K = 8 with FP32 + FP64 rescoring of the raw FP32 data.
    Returns a dict of device tensors: index (N_i,) int32 and diagnostics.
    """
    if K is None:
        K = 1 if precision == "fp32" else 8
    if rescore_fp64 is None:
        rescore_fp64 = precision != "fp32"
    if images.N_p != slices.N_p:
        raise ValueError(f"images have {images.N_p} pixels, slices {slices.N_p}")
    if alpha is None:
        alpha = scaling_ratio(images.row_sum, images.row_sumsq, slices.row_sum, slices.row_sumsq,
                              images.N_p, stream)
    with _span(timers, "residual"):
        S = slices.operand(precision, center, stream)
        D = images.operand(precision, alpha, slices.center if center else None, stream)
    one = _one(alpha.device)
    idx, val = topk(precision, S, S.sumsq, slices.N_o, D, images.N_i, images.N_p, one,
                    K, workspace, stream, timers)
    out = dict(topk_idx=idx, topk_val=val, alpha=alpha)
    if rescore_fp64:
        order = None
        if slices.rank is not None:
            # images whose best candidates are neighbours in orientation space
            # go to neighbouring workgroups, so their candidate slices are
            # re-read from cache rather than from HBM
            with _span(timers, "order"):
                key = slices.rank[idx[:, 0].long().clamp(0, slices.N_o - 1)]
                order = torch.argsort(key).to(torch.int32)
        r = rescore(slices.f32, images.f32, images.N_p, alpha, idx, stream, timers,
                    image_order=order)
        out.update(index=r["best"], best_d64=r["best_d64"], d32=r["d32"], d64=r["d64"])
    else:
        out["index"] = idx[:, 0]
    return out
