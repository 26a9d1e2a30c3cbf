"""Batching form: uneven image batches matched against a hierarchy of
orientation grids, coarse to fine, entirely on the device.

Levels are near-uniform quaternion grids of increasing size (1k, 10k, 100k
by default).  Each node of level l + 1 takes as parents its ``n_parents``
nearest nodes of level l (largest |q . q'|, found with the FP32 top-K
kernel on zero-padded quaternions and their negatives); the children of a
level-l node are the nodes that have it among their parents.  Nodes are
stored in an order where siblings are adjacent and neighbouring parents are
near in index (level 0 in rotation-vector Morton order, finer levels by
parent), so the slices that concurrent tasks read lie close together.
Children also include those of the node nearest each node's twin (the half
turn about the beam axis after it): for a centrosymmetric intensity and a
nearly flat detector cap the two give almost the same slice.

A pass (``HierarchicalMatcher.run``):
  * per batch (request) scaling by its own std ratio against level 0's
    slices (``mtip_segment_scaling_ratio``), images centred by level 0's
    mean slice and cast to bf16;
  * level 0 (stored in Morton order) split into ``parts`` contiguous ranges,
    i.e. regions of orientation space, and a dense top-``k0`` against each
    (the MFMA top-K kernel): the beam holds the best nodes of every region,
    not only those of the best basin.  Slices of pseudo-symmetric
    orientations (the twin half turn, and others of a nearly centrosymmetric
    pattern) can rank a wrong basin's coarse nodes above the true basin's;
  * per finer level, (image, winner) entries grouped by winner
    (``mtip_group_by_winner``) and a persistent grouped top-K over the
    children of each winner (``mtip_group_topk``).  A middle level keeps each
    entry's top-``kl_mid`` children as the next level's winners, repeats
    removed (``mtip_unique_rows``; no merge: the beam stays spread over the
    basins it reached);
    the last level merges every entry's top-K into the image's K best
    distinct candidates (``mtip_merge_distinct``);
  * FP32 then FP64 rescoring of the last level's top-K (``mtip_rescore``).
No step needs the host, so a pass over fixed batch sizes is captured once as
a HIP graph (``capture``) and replayed (``replay``).

Measured at 4 x 32 x 128 px, 1k / 10k / 100k, 4,087 Poisson images, against
the flat BF16 + FP64 search over the 100k level: one part with the 4 best
(and 4 distinct candidates kept per middle level) agreed for 97.3 % (M = 81)
and 91.4 % (M = 121) of the images, its misses 0.4-1.4 % worse in distance
and mostly in another basin (median 161 degrees away); 2 parts x 8 with 4
children per entry agree for 99.7 % and 99.1 %.

The reference matches every image against one flat grid
(spinifel/sequential/orientation_matching.py:189-256); this form is an
addition beside it, and its result is checked against the flat search over
the finest level.
"""
import numpy as np
import torch

from . import _lib, engine
from .engine import _ptr, _stream
from .quaternions import _rotvec, morton_order, uniform_quaternions  # noqa: F401


def children_csr(parents, n_coarse):
    """CSR (offsets, child indices) of the coarse nodes from each fine node's parents.

    parents: (N_fine, k) int array of coarse indices.  Children of a node are
    in ascending fine index.
    """
    parents = np.asarray(parents, dtype=np.int64)
    n_fine, k = parents.shape
    p = parents.ravel()
    c = np.repeat(np.arange(n_fine, dtype=np.int64), k)
    order = np.lexsort((c, p))
    counts = np.bincount(p, minlength=n_coarse)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return off, c[order].astype(np.int32)


def nearest_parents(coarse, fine, k, device):
    """Indices (N_fine, k) of each fine quaternion's k nearest coarse ones (GPU).

    Nearest by |q_c . q_f| (q and -q are one rotation): the coarse set enters
    with both signs, as rows of 64 FP32 columns (zero padded), so the FP32
    top-K of ||s||^2 - 2 s . d = 1 - 2 q_c . q_f picks the largest dots.
    """
    dev = torch.device("cuda", device)
    nc, nf = len(coarse), len(fine)
    A = torch.zeros((2 * nc, 64), dtype=torch.float32, device=dev)
    cq = torch.as_tensor(np.asarray(coarse, np.float32)).to(dev)
    A[:nc, :4] = cq
    A[nc:, :4] = -cq
    B = torch.zeros((nf, 64), dtype=torch.float32, device=dev)
    B[:, :4] = torch.as_tensor(np.asarray(fine, np.float32)).to(dev)
    ones = torch.ones(2 * nc, dtype=torch.float64, device=dev)
    S, D = engine.Operand(A, None, ones), engine.Operand(B)
    idx, _ = engine.topk("fp32", S, ones, 2 * nc, D, nf, 64, engine._one(dev), k)
    return (idx.cpu().numpy().astype(np.int64)) % nc


def quat_mul(a, b):
    """Hamilton products a * b of (n, 4) [w, x, y, z] arrays."""
    w1, x1, y1, z1 = np.asarray(a, np.float64).T
    w2, x2, y2, z2 = np.asarray(b, np.float64).T
    return np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2],
                    axis=1)


def twin_quaternions(q, axis=(0.0, 0.0, 1.0)):
    """q's twins: the half turn about the beam axis applied after q.

    The model's intensity is centrosymmetric (a real autocorrelation) and a
    detector near the beam samples an almost flat cap, so the slice at R
    and at Rz(pi) R nearly coincide; refinement follows both branches."""
    ax = np.asarray(axis, np.float64)
    ax = ax / np.linalg.norm(ax)
    h = np.tile(np.concatenate([[0.0], ax]), (len(q), 1))       # [cos(pi/2), sin(pi/2) axis]
    t = quat_mul(h, q)
    t[t[:, 0] < 0] *= -1.0
    return t


def union_csr(off, ch, partner):
    """CSR whose row c is the union of rows c and partner[c] (ascending, distinct)."""
    rows = [np.union1d(ch[off[c]:off[c + 1]], ch[off[p]:off[p + 1]]) for c, p in enumerate(partner)]
    o = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    return o, (np.concatenate(rows) if rows else np.zeros(0)).astype(np.int32)


class Hierarchy:
    """Orientation grids of increasing size and their parent / child links.

    twins: each coarse node's children also include those of the node
    nearest its twin (``twin_quaternions`` about ``beam_axis``)."""

    def __init__(self, sizes=(1000, 10000, 100000), n_parents=4, device=0, levels=None,
                 twins=True, beam_axis=(0.0, 0.0, 1.0)):
        if n_parents not in (1, 2, 4, 8):
            raise ValueError("n_parents must be 1, 2, 4 or 8")
        engine.require_gpu()
        raw = [np.asarray(q, np.float64) for q in levels] if levels is not None else \
            [uniform_quaternions(n) for n in sizes]
        if len(raw) < 1:
            raise ValueError("need at least one level")
        self.device = torch.device("cuda", device)
        self.n_parents = n_parents
        q0 = raw[0][morton_order(raw[0])]
        self.quats = [q0]
        self.children = []          # per transition: (offsets, indices) numpy int32
        for q in raw[1:]:
            prev = self.quats[-1]
            par = nearest_parents(prev, q, n_parents, device)
            order = np.lexsort((morton_order(q).argsort(), par[:, 0]))
            q, par = q[order], par[order]
            self.quats.append(q)
            off, ch = children_csr(par, len(prev))
            if twins:
                partner = nearest_parents(prev, twin_quaternions(prev, beam_axis), 1, device)[:, 0]
                off, ch = union_csr(off, ch, partner)
            self.children.append((off, ch))
        self.sizes = [len(q) for q in self.quats]
        dev = self.device
        self.quats_dev = [torch.as_tensor(q).to(dev) for q in self.quats]
        self.children_dev = [(torch.as_tensor(o).to(dev), torch.as_tensor(c).to(dev))
                             for o, c in self.children]

    @property
    def n_levels(self):
        return len(self.quats)


class HierarchyModel:
    """One model's slices at every level of a hierarchy, resident in HBM:
    FP32 rows for rescoring and bf16 residuals (centred by level 0's mean
    slice) for the top-K passes."""

    def __init__(self, hierarchy, plan, q, scale, order=None, stream=None, timers=None):
        self.h = hierarchy
        self.levels = []
        for hq in hierarchy.quats_dev:
            sl = engine.SliceSet.compute(plan, hq, q, scale, stream=stream, timers=timers,
                                         order=order)
            self.levels.append(sl)
        self.N_p = self.levels[0].N_p
        self.center = self.levels[0].center
        self.ops = [engine.residual(sl.f32, "bf16", sl.N_p, center=self.center, stream=stream)
                    for sl in self.levels]
        self._ops8 = {}
        self.bad = sum(int(sl.bad.item()) for sl in self.levels)

    def op8(self, level, stream=None):
        """The level's centred residual as e4m3 rows with per-row scales (built once)."""
        if level not in self._ops8:
            sl = self.levels[level]
            self._ops8[level] = engine.residual(sl.f32, "fp8", sl.N_p, center=self.center,
                                                stream=stream)
        return self._ops8[level]


class HierarchicalMatcher:
    """Refinement through a HierarchyModel's levels for a set of image batches.

    parts / k0: level 0 is searched as ``parts`` contiguous (Morton) ranges
    with the top ``k0`` of each; kl_mid: children kept per entry on middle
    levels; K: candidates of the last level rescored in FP32 / FP64;
    fp8_levels: grouped levels (1 .. n_levels - 1) whose cross terms run on
    e4m3 operands with per-row scales instead of bf16 (the images' e4m3
    residual is built once per pass).  At 4 x 32 x 128 px and 1k / 10k /
    100k levels, e4m3 on the 10k level left the agreement with the flat
    search within 0.03 % (M = 81 and 121); on the 100k level it fell from
    99.1 to 97.1 % at M = 121, so the default keeps bf16 everywhere."""

    def __init__(self, model, K=8, parts=2, k0=8, kl_mid=4, fp8_levels=()):
        if K not in (1, 2, 4, 8) or k0 not in (1, 2, 4, 8) or kl_mid not in (1, 2, 4, 8):
            raise ValueError("K, k0 and kl_mid must be 1, 2, 4 or 8")
        fp8_levels = tuple(sorted(set(int(l) for l in fp8_levels)))
        if any(l < 1 or l >= model.h.n_levels for l in fp8_levels):
            raise ValueError("fp8_levels must be grouped levels 1 .. n_levels - 1")
        self.fp8_levels = fp8_levels
        n0 = model.levels[0].N_o
        if parts < 1 or n0 // parts < k0:
            raise ValueError(f"{parts} parts of {n0} level-0 orientations cannot each give {k0}")
        self.m = model
        self.K = K
        self.parts = int(parts)
        self.k0 = int(k0)
        self.kl_mid = int(kl_mid)
        self.graph = None
        self._static = None

    def _pass(self, images, offsets, stream=None, timers=None, alpha=None):
        lib = _lib.load()
        m, h = self.m, self.m.h
        dev = images.f32.device
        N_i, N_p = images.N_i, images.N_p
        L0 = m.levels[0]
        st = _stream(stream)
        with engine._span(timers, "scale"):
            if alpha is None:
                alpha = engine.segment_scaling_ratio(images.row_sum, images.row_sumsq, offsets,
                                                     L0.row_sum, L0.row_sumsq, N_p, stream)
            D = images.operand("bf16", alpha, m.center, stream)
            D8 = images.operand("fp8", alpha, m.center, stream) if self.fp8_levels else None
        last = h.n_levels - 1
        one = engine._one(dev)
        op0 = m.ops[0]
        n0 = L0.N_o
        k0 = self.K if last == 0 else self.k0
        with engine._span(timers, "level0"):
            cols = []
            for p in range(self.parts):
                a, b = p * n0 // self.parts, (p + 1) * n0 // self.parts
                sub = engine.Operand(op0.data[a:b], None, op0.sumsq[a:b])
                ix, vl = engine.topk("bf16", sub, sub.sumsq, b - a, D, N_i, N_p, one, k0, None,
                                     stream)
                cols.append((ix + a, vl))
        if self.parts == 1:
            idx, val = cols[0]
        else:
            idx = torch.cat([c[0] for c in cols], dim=1).contiguous()
            val = torch.cat([c[1] for c in cols], dim=1).contiguous()
        if last == 0:
            if self.parts > 1:
                merged_i = torch.empty((N_i, self.K), dtype=torch.int32, device=dev)
                merged_v = torch.empty((N_i, self.K), dtype=torch.float32, device=dev)
                _lib.check(lib.mtip_merge_distinct(_ptr(idx), _ptr(val), N_i, idx.shape[1],
                                                   self.K, _ptr(merged_i), _ptr(merged_v), st),
                           "mtip_merge_distinct")
                idx = merged_i
        levels = [idx]
        ld = D.data.shape[1]
        for l in range(1, h.n_levels):
            T = idx.shape[1]
            n_groups = h.sizes[l - 1]
            kl = self.K if l == last else self.kl_mid
            nb = int(lib.mtip_group_workspace_size(N_i * T, n_groups))
            ws = torch.empty(nb, dtype=torch.uint8, device=dev)
            with engine._span(timers, f"group{l}"):
                _lib.check(lib.mtip_group_by_winner(_ptr(idx), T, T, N_i, n_groups,
                                                    _ptr(ws), nb, st), "mtip_group_by_winner")
            eidx = torch.full((N_i * T, kl), -1, dtype=torch.int32, device=dev)
            evals = torch.full((N_i * T, kl), float("inf"), dtype=torch.float32, device=dev)
            off, ch = h.children_dev[l - 1]
            with engine._span(timers, f"level{l}"):
                if l in self.fp8_levels:
                    op = m.op8(l, stream)
                    ld8 = D8.data.shape[1]
                    _lib.check(lib.mtip_group_topk_prec(
                        _lib.MTIP_FP8, _ptr(ws), T, N_i, n_groups, _ptr(off), _ptr(ch),
                        _ptr(op.data), op.data.shape[1], _ptr(op.scale), _ptr(op.sumsq),
                        _ptr(D8.data), ld8, _ptr(D8.scale), ld8, kl, _ptr(eidx), _ptr(evals),
                        dev.index or 0, st), "mtip_group_topk_prec")
                else:
                    op = m.ops[l]
                    _lib.check(lib.mtip_group_topk(_ptr(ws), T, N_i, n_groups, _ptr(off),
                                                   _ptr(ch), _ptr(op.data), op.data.shape[1],
                                                   _ptr(op.sumsq), _ptr(D.data), ld, ld, kl,
                                                   _ptr(eidx), _ptr(evals), dev.index or 0, st),
                               "mtip_group_topk")
            if l == last:
                idx = torch.empty((N_i, kl), dtype=torch.int32, device=dev)
                val = torch.empty((N_i, kl), dtype=torch.float32, device=dev)
                with engine._span(timers, f"merge{l}"):
                    _lib.check(lib.mtip_merge_distinct(_ptr(eidx), _ptr(evals), N_i, T * kl, kl,
                                                       _ptr(idx), _ptr(val), st),
                               "mtip_merge_distinct")
            else:
                # every entry's top-kl children become winners, a child reached
                # from two entries once (repeats become -1 and are left out)
                idx = eidx.reshape(N_i, T * kl)
                with engine._span(timers, f"unique{l}"):
                    _lib.check(lib.mtip_unique_rows(_ptr(idx), N_i, T * kl, st),
                               "mtip_unique_rows")
            levels.append(idx)
        # rescore with the images grouped by their best candidate, so that
        # neighbouring workgroups read the same candidate slices
        n_last = h.sizes[last]
        nb = int(lib.mtip_group_workspace_size(N_i, n_last))
        ws = torch.empty(nb, dtype=torch.uint8, device=dev)
        with engine._span(timers, "order"):
            _lib.check(lib.mtip_group_by_winner(_ptr(idx), idx.shape[1], 1, N_i, n_last, _ptr(ws),
                                                nb, st), "mtip_group_by_winner")
        po = int(lib.mtip_group_perm_offset(N_i, n_last))
        order = ws[po:po + 4 * N_i].view(torch.int32)
        r = engine.rescore(m.levels[last].f32, images.f32, N_p, alpha, idx, stream, timers,
                           image_order=order)
        return dict(index=r["best"], best_d64=r["best_d64"], topk_idx=idx, alpha=alpha,
                    levels=levels)

    def run(self, images, offsets, stream=None, timers=None, alpha=None):
        """One eager pass.  images: engine.ImageSet of all batches concatenated;
        offsets: batch boundaries (n_batches + 1); alpha: optional per-image
        FP64 device scale replacing the per-batch std ratio.  Returns device
        tensors: index (N_i,) int32 into the last level, best_d64, topk_idx,
        alpha, levels (each level's top candidates)."""
        return self._pass(images, offsets, stream, timers, alpha)

    def capture(self, images, offsets, warmup=1):
        """Capture one pass over these (resident) images as a HIP graph."""
        offsets = torch.as_tensor(np.asarray(offsets, np.int64)).to(images.f32.device)
        s = torch.cuda.Stream(device=images.f32.device)
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self._pass(images, offsets, s)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            out = self._pass(images, offsets)
        self.graph, self._static = g, out
        return out

    def replay(self):
        """Replay the captured pass; returns the captured output tensors."""
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()
        return self._static

    def orientations(self, index):
        """Quaternions (N, 4) of last-level indices (host numpy)."""
        ix = np.asarray(index.cpu().numpy() if torch.is_tensor(index) else index, np.int64)
        return self.m.h.quats[-1][ix]


def log_uniform_batches(n_batches, lo, hi, rng):
    """Batch sizes floor(exp(U(ln lo, ln hi)))."""
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), int(n_batches)))).astype(np.int64)
