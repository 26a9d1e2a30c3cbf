"""Merge (solve_ac) on the GPU, drop-in for spinifel's ``Merge``
(spinifel/sequential/autocorrelation.py:35-225):

    merge = Merge(settings, slices_, pixel_position_reciprocal,
                  pixel_distance_reciprocal, nufft)
    ac = merge.solve_ac(generation, orientations, ac_estimate)

The step right after orientation matching: with the images' orientations
known, solve (A^H A + rlambda I + flambda F^H P F) ac = A^H b + rlambda x0
for the autocorrelation by conjugate gradients, A^H A applied as a
convolution on the 2M grid whose kernel is the type-1 transform of ones
(the reference's "convolution technique", :124-166).

Where the work runs:
  * both type-1 transforms (the convolution kernel at M_ups of ones, the data
    term at M with the support) -- libmtip's spreading kernel and transform
    passes, positions formed on the fly from the quaternions (no H, K, L
    arrays), the images resident in HBM as FP32;
  * the per-iteration FFTs (the convolution on the M_ups^3 grid and the
    Fourier regulariser on M^3) -- rocFFT through torch.fft, in FP64, as
    real-to-complex transforms: every operand is real and every product
    Hermitian, so irfftn(rfftn(.) * half-spectrum) equals the reference's
    real part of ifftn(fftn(.) * spectrum) (the regulariser's mask is
    replaced by its even part (m(k) + m(-k)) / 2, which is what the real
    part keeps);
  * the CG iteration -- FP64 vector updates on the device, one host sync
    per iteration for the stopping test (the reference's cupyx cg syncs the
    same way).
Differences a caller can observe:
  * orientations=None draws uniform random quaternions from numpy's
    generator (the reference calls skopi.get_random_quat, absent here);
  * the transform tolerance is ``settings.nufft_eps`` (1e-6 by default;
    the reference fixes 1e-12), the spreading sums FP32 weights into an FP64
    grid; CG runs in FP64 whatever ``use_single_prec`` says;
  * ``solve_ac`` returns the real float64 (M, M, M) array like the
    reference; ``last`` holds the device result and the iteration count.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib, engine, filters
from .autocorrelation import get_plan, gen_nonuniform_positions
from .quaternions import random_quaternions


def antisupport(M, oversampling):
    """F_antisupport: |q| > pi / oversampling on linspace(-pi, pi, M)^3
    (sequential/autocorrelation.py:75-82), with the reference's symmetry asserts."""
    lu = np.linspace(-np.pi, np.pi, M)
    Hu_, Ku_, Lu_ = np.meshgrid(lu, lu, lu, indexing="ij")
    F = np.sqrt(Hu_ ** 2 + Ku_ ** 2 + Lu_ ** 2) > np.pi / oversampling
    assert np.all(F == F[::-1, :, :])
    assert np.all(F == F[:, ::-1, :])
    assert np.all(F == F[:, :, ::-1])
    assert np.all(F == F[::-1, ::-1, ::-1])
    return F


def gaussian_filter(a, sigma=0.5, truncate=4.0):
    """scipy.ndimage.gaussian_filter(a, sigma) on a device tensor (FP64,
    separable, mode 'reflect'; spinifel/sequential/autocorrelation.py:201)."""
    return filters.gaussian_filter(a, sigma, "reflect", truncate)


def cg(matvec, b, x0, maxiter, tol=1e-5, callback=None):
    """Conjugate gradients on device vectors as cupyx/scipy
    ``cg(A, b, x0=x0, tol=tol, maxiter=maxiter, callback=callback)`` with no
    preconditioner (spinifel/sequential/autocorrelation.py:211-212): stop
    after the update that brings ||r|| <= tol ||b||, or after maxiter
    updates.  Returns (x, info, iterations), info = iterations when not
    converged."""
    x = x0.clone()
    b_norm = float(torch.linalg.norm(b))
    if b_norm == 0.0:
        return b.clone(), 0, 0
    atol = tol * b_norm
    r = b - matvec(x)
    rho = None
    p = None
    it = 0
    resid = float(torch.linalg.norm(r))
    while it < maxiter:
        rho1, rho = rho, torch.dot(r, r)
        p = r.clone() if it == 0 else r + (rho / rho1) * p
        q = matvec(p)
        alpha = rho / torch.dot(p, q)
        x += alpha * p
        r -= alpha * q
        it += 1
        if callback is not None:
            callback(x)
        resid = float(torch.linalg.norm(r))
        if resid <= atol:
            break
    info = it if (it == maxiter and not resid <= atol) else 0
    return x, info, it


def cg_fused(matvec, b, x0, maxiter, tol=1e-5, callback=None):
    """The same iteration as ``cg`` with libmtip's vector kernels: the step
    lengths stay on the device (rho = r.r of the previous update serves as the
    next iteration's rho) and dot products reduce in a fixed order.  The
    stopping test after update k reads ||r_k|| through pinned memory while
    iteration k + 1's direction and operator application already run, so the
    GPU does not idle on the host's check; on convergence that look-ahead
    (which changes neither x nor r) is dropped.  Returns (x, info, iterations)."""
    lib = _lib.load()
    dev = b.device
    st = engine._stream()
    P = engine._ptr
    n = b.numel()
    work = torch.empty(int(lib.mtip_dot_workspace_size()) // 8, dtype=torch.float64, device=dev)
    wbytes = work.numel() * 8
    sc = torch.zeros(4, dtype=torch.float64, device=dev)    # rho[0], rho[1], pq, b.b
    host = torch.empty(1, dtype=torch.float64, pin_memory=True)
    ready = torch.cuda.Event()

    def dot(a, c, slot):
        _lib.check(lib.mtip_dot(P(a), P(c), n, P(work), wbytes,
                                ctypes.c_void_p(sc.data_ptr() + 8 * slot), st), "mtip_dot")

    def ptr(slot):
        return ctypes.c_void_p(sc.data_ptr() + 8 * slot)

    x = x0.clone().contiguous()
    b = b.contiguous()
    dot(b, b, 3)
    b_norm = float(torch.sqrt(sc[3]))
    if b_norm == 0.0:
        return b.clone(), 0, 0
    atol = tol * b_norm
    r = (b - matvec(x)).contiguous()
    p = torch.empty_like(r)
    it, cur = 0, 0
    dot(r, r, cur)
    resid = float(torch.sqrt(sc[cur]))
    pending = False          # ||r|| of the last update is on its way to `host`
    while True:
        if it >= maxiter:
            if pending:
                ready.synchronize()
                resid = math.sqrt(float(host[0]))
            break
        _lib.check(lib.mtip_cg_direction(P(p), P(r), ptr(cur), ptr(1 - cur), n, int(it == 0), st),
                   "mtip_cg_direction")
        q = matvec(p).contiguous()
        dot(p, q, 2)
        if pending:
            ready.synchronize()
            pending = False
            resid = math.sqrt(float(host[0]))
            if resid <= atol:
                break
        _lib.check(lib.mtip_cg_update(P(x), P(r), P(p), P(q), ptr(cur), ptr(2), n, st),
                   "mtip_cg_update")
        it += 1
        if callback is not None:
            callback(x)
        cur = 1 - cur
        dot(r, r, cur)
        host.copy_(sc[cur:cur + 1], non_blocking=True)
        ready.record()
        pending = True
    info = it if (it == maxiter and not resid <= atol) else 0
    return x, info, it


class Merge:
    """spinifel/sequential/autocorrelation.py:35-225 on the GPU."""

    def __init__(self, settings, slices_, pixel_position_reciprocal, pixel_distance_reciprocal,
                 nufft=None, eps=None):
        engine.require_gpu()
        self.M = int(settings.M)
        self.M_ups = int(getattr(settings, "M_ups", 2 * self.M))
        self.device = torch.device("cuda", int(getattr(settings, "device", 0)))
        data = slices_ if torch.is_tensor(slices_) else np.asarray(slices_)
        self.N_images = int(data.shape[0])
        self.N = int(np.prod(tuple(data.shape)))
        self.reciprocal_extent = float(np.asarray(pixel_distance_reciprocal).max())
        self.use_reciprocal_symmetry = True
        self.maxiter = int(getattr(settings, "solve_ac_maxiter", 100))
        self.oversampling = settings.oversampling
        self.eps = float(getattr(settings, "nufft_eps", 1e-6) if eps is None else eps)
        self.sign = 1
        self.pixel_position_reciprocal = np.asarray(pixel_position_reciprocal, np.float64)
        self.det_shape = tuple(self.pixel_position_reciprocal.shape[1:])
        self.N_pixels = int(np.prod(self.det_shape))
        self.mult = np.pi / (self.reciprocal_extent * self.oversampling)
        self.rlambda = 1 / self.N / 1000
        self.flambda = 1e3
        # nuvect_Db = data * weights with weights = 1 (:66-69), resident as FP32
        self._data = torch.as_tensor(data).reshape(self.N_images, -1).to(
            self.device, torch.float32).contiguous()
        if self._data.shape[1] != self.N_pixels:
            raise ValueError("slices_ do not match pixel_position_reciprocal")
        self._q = torch.as_tensor(np.ascontiguousarray(
            self.pixel_position_reciprocal.reshape(3, -1))).to(self.device)
        M = self.M
        F = antisupport(M, self.oversampling)
        self.F_antisupport = torch.as_tensor(F, device=self.device)
        m = np.fft.ifftshift(F).astype(np.float64)
        neg = (-np.arange(M)) % M
        m_even = 0.5 * (m + m[np.ix_(neg, neg, neg)])
        self._reg_mask = torch.as_tensor(np.ascontiguousarray(m_even[:, :, :M // 2 + 1]),
                                         device=self.device)
        self.nufft = nufft
        self._plan = get_plan(self.M, self.eps, self.device.index or 0)
        self._plan_ups = get_plan(self.M_ups, self.eps, self.device.index or 0)
        # zero-padded operand of the convolution; only [:M]^3 is ever written
        self._ups = torch.zeros((self.M_ups,) * 3, dtype=torch.float64, device=self.device)
        self._us = torch.empty((self.M,) * 3, dtype=torch.float64, device=self.device)

        def callback(xk):
            callback.counter += 1

        self.callback = callback
        self.callback.counter = 0
        self.last = None

    # ------------------------------------------------------------ operators

    def get_non_uniform_positions(self, orientations):
        """H, K, L (N_images,) + det_shape of inv(R) q (:85-108); the solve
        itself forms them inside the spreading kernel."""
        return gen_nonuniform_positions(orientations, self.pixel_position_reciprocal)

    # The reference wraps both per-iteration transforms in ifftshift / fftshift
    # of the real-space grid.  A product in frequency space is a circular
    # convolution, which commutes with cyclic shifts, and fftshift undoes
    # ifftshift, so fftshift(ifft(fft(ifftshift(x)) S)) = ifft(fft(x) S):
    # the two rolls per transform are dropped (the kernel's own ifftshift,
    # applied once in setup_linops, stays).

    def fourier_reg(self, uvect, support):
        """:110-122 -- support, FFT, antisupport mask, inverse FFT, support."""
        M = self.M
        g = uvect.view(M, M, M) * support
        X = torch.fft.rfftn(g)
        X *= self._reg_mask
        reg = torch.fft.irfftn(X, s=(M, M, M))
        return (reg * support).reshape(-1)

    def core_problem_convolution(self, uvect, F_ugrid_conv_half, ac_support):
        """:124-145 -- zero-pad to M_ups at [:M], multiply by the kernel's
        spectrum, crop [:M], support."""
        M, Mu = self.M, self.M_ups
        ups = self._ups
        ups[:M, :M, :M] = uvect.view(M, M, M) * ac_support
        X = torch.fft.rfftn(ups)
        X *= F_ugrid_conv_half
        out = torch.fft.irfftn(X, s=(Mu,) * 3)[:M, :M, :M] * ac_support
        return out.reshape(-1)

    def _quats(self, orientations):
        q = torch.as_tensor(np.ascontiguousarray(orientations, np.float64).reshape(-1, 4)) \
            if not torch.is_tensor(orientations) else orientations
        q = q.to(self.device, torch.float64).contiguous()
        if q.shape[0] != self.N_images:
            raise ValueError(f"{q.shape[0]} orientations for {self.N_images} images")
        return q

    def _points(self, quats):
        """Every (image, pixel) point, rotated and scaled (radians), in the
        Morton order of 4-cell cubes of the M grid, and the subset whose data
        are non-zero (a zero value adds nothing to the data term).  Spreading
        runs of 128 spatially close points keeps each tile's LDS box small,
        whichever images they come from: at M = 81, 1000 images x 16,384 px
        this path (order 2.7 ms, data 0.8 ms, kernel at M_ups 6.3 ms) replaces
        per-image detector tiles (3.9 and 12.0 ms)."""
        x, y, z = engine.rotated_points(quats, self._q, self.mult)
        perm = engine.spatial_order(x, y, z, 4 * 2 * np.pi / self._plan.n_fine)
        x, y, z = x[perm], y[perm], z[perm]
        v = self._data.reshape(-1)[perm]
        nz = torch.nonzero(v).reshape(-1)
        return (x, y, z), (x[nz], y[nz], z[nz], v[nz])

    def setup_linops(self, orientations, ac_support, x0):
        """W (a callable on FP64 device vectors) and d (:147-190)."""
        quats = self._quats(orientations)
        (x, y, z), data_pts = self._points(quats)
        ugrid_conv = self._plan_ups.adjoint_points(x, y, z, values=None,
                                                   out_scale=1.0 / self.M_ups ** 3, spatial=False)
        del x, y, z
        F_half = torch.fft.rfftn(torch.fft.ifftshift(ugrid_conv)) / self.M ** 3

        lib = _lib.load()
        P = engine._ptr
        sup = ac_support.contiguous()

        def W_matvec(uvect):
            # the same operator as core_problem_convolution + rlambda u +
            # flambda fourier_reg, with the padding and the recombination in
            # two kernels around the four FFTs
            M, Mu = self.M, self.M_ups
            u = uvect.contiguous()
            st = engine._stream()
            _lib.check(lib.mtip_merge_pad(P(u), P(sup), M, Mu, P(self._ups), P(self._us), st),
                       "mtip_merge_pad")
            X = torch.fft.rfftn(self._ups)
            X *= F_half
            y_ups = torch.fft.irfftn(X, s=(Mu,) * 3).contiguous()
            R = torch.fft.rfftn(self._us)
            R *= self._reg_mask
            y_reg = torch.fft.irfftn(R, s=(M,) * 3).contiguous()
            out = torch.empty(M ** 3, dtype=torch.float64, device=self.device)
            _lib.check(lib.mtip_merge_combine(P(y_ups), P(y_reg), P(u), P(sup), M, Mu,
                                              float(self.rlambda), float(self.flambda), P(out),
                                              st), "mtip_merge_combine")
            return out

        xd, yd, zd, vd = data_pts
        uvect_ADb = self._plan.adjoint_points(xd, yd, zd, values=vd, support=ac_support,
                                              out_scale=1.0 / self.M ** 3,
                                              spatial=False).reshape(-1)
        d = uvect_ADb + self.rlambda * x0
        return W_matvec, d

    # ------------------------------------------------------------ solve

    def solve_ac(self, generation, orientations=None, ac_estimate=None):
        """:192-225 -> ac, real float64 (M, M, M)."""
        M = self.M
        if orientations is None:
            orientations = random_quaternions(self.N_images, np.random.default_rng())
        if ac_estimate is None:
            ac_support = torch.ones((M,) * 3, dtype=torch.float64, device=self.device)
            ac_est = torch.zeros((M,) * 3, dtype=torch.float64, device=self.device)
        else:
            ac_est = torch.as_tensor(ac_estimate).to(self.device, torch.float64)
            ac_smoothed = gaussian_filter(ac_est, 0.5)
            ac_support = (ac_smoothed > 1e-12).to(torch.float64)
            ac_est = ac_est * ac_support
            if not torch.is_tensor(ac_estimate):
                # the reference scales the caller's array in place (:203)
                ac_estimate *= ac_support.cpu().numpy().astype(ac_estimate.dtype)
        x0 = ac_est.reshape(-1)
        W, d = self.setup_linops(orientations, ac_support, x0)
        before = self.callback.counter
        ret, info, it = cg_fused(W, d, x0, self.maxiter, callback=self.callback)
        ac = ret.view(M, M, M)
        self.last = dict(ac=ac, info=info, iterations=self.callback.counter - before,
                         support=ac_support, d=d, W=W)
        return ac.cpu().numpy()
