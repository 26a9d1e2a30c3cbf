"""ORACLE -- test infrastructure, not product code.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module.

FP64 CPU restatement of spinifel's merge step, ``Merge.solve_ac``
(spinifel/sequential/autocorrelation.py:35-225), with numpy FFTs:

* ``antisupport``: ``F_antisupport`` of :75-82 (``|q| > pi / oversampling`` on
  ``linspace(-pi, pi, M)``);
* ``fourier_reg``: :110-122 (support, FFT, mask, inverse, support, real);
* ``core_problem_convolution``: :124-145 (zero-pad to M_ups at [:M], FFT
  product with the kernel's transform, crop [:M], support, real);
* ``conv_kernel``: :153-166 (``adjoint`` of ones at M_ups, FFT / M^3);
* ``data_term``: :183-191 (``adjoint`` of the data with the support at M,
  plus ``rlambda x0``);
* ``cg``: the conjugate-gradient iteration of cupyx.scipy.sparse.linalg.cg /
  scipy.sparse.linalg.cg called at :211-212 (identity preconditioner,
  stopping ``||r|| <= tol ||b||`` after each update, tol = 1e-5, at most
  ``maxiter`` updates, callback per update);
* ``gaussian_filter``: scipy.ndimage.gaussian_filter(a, 0.5) of :201
  (truncate 4.0 -> radius 2, mode 'reflect' = half-sample symmetric);
* ``solve_ac``: :192-225 with ``orientations`` given (the reference draws
  ``skp.get_random_quat`` when they are None; skopi is absent).

Parity status: ``core_problem_convolution`` is pinned against the brute-force
A^H A (type-2 then type-1 direct sums) it stands for, with its 1 / M_ups^3
factor; ``fourier_reg`` by projection properties (idempotent with support 1,
mask of ones = identity, zeros = 0) and a delta closed form; ``cg`` against
scipy.sparse.linalg.cg and numpy.linalg.solve on SPD systems;
``gaussian_filter`` against scipy.ndimage (importable here) -- see
tests/test_oracle_merge.py.
"""
import numpy as np

from . import es_nufft


def antisupport(M, oversampling):
    """F_antisupport (sequential/autocorrelation.py:75-82)."""
    lu = np.linspace(-np.pi, np.pi, M)
    Hu_, Ku_, Lu_ = np.meshgrid(lu, lu, lu, indexing="ij")
    Qu_ = np.sqrt(Hu_ ** 2 + Ku_ ** 2 + Lu_ ** 2)
    return Qu_ > np.pi / oversampling


def fourier_reg(uvect, support, F_antisupport, M, use_recip_sym=True):
    """sequential/autocorrelation.py:110-122."""
    ugrid = uvect.reshape((M,) * 3) * support
    F_ugrid = np.fft.fftn(np.fft.ifftshift(ugrid))
    F_reg = F_ugrid * np.fft.ifftshift(F_antisupport)
    reg = np.fft.fftshift(np.fft.ifftn(F_reg))
    uvect = (reg * support).reshape(-1)
    return uvect.real if use_recip_sym else uvect


def core_problem_convolution(uvect, M, F_ugrid_conv_, M_ups, ac_support, use_recip_sym=True):
    """sequential/autocorrelation.py:124-145."""
    ugrid = uvect.reshape((M,) * 3) * ac_support
    ugrid_ups = np.zeros((M_ups,) * 3, dtype=np.result_type(uvect.dtype, np.float64))
    ugrid_ups[:M, :M, :M] = ugrid
    F_ugrid_ups = np.fft.fftn(np.fft.ifftshift(ugrid_ups))
    F_ugrid_conv_out_ups = F_ugrid_ups * F_ugrid_conv_
    ugrid_conv_out_ups = np.fft.fftshift(np.fft.ifftn(F_ugrid_conv_out_ups))
    ugrid_conv_out = ugrid_conv_out_ups[:M, :M, :M] * ac_support
    if use_recip_sym:
        ugrid_conv_out = ugrid_conv_out.real
    return ugrid_conv_out.reshape(-1)


def conv_kernel(H_, K_, L_, M, M_ups, eps=1e-12):
    """ugrid_conv and F_ugrid_conv_ (sequential/autocorrelation.py:153-166)."""
    ugrid_conv = es_nufft.adjoint(np.ones(np.size(H_)), H_, K_, L_, M_ups, True, 1, eps)
    return ugrid_conv, np.fft.fftn(np.fft.ifftshift(ugrid_conv)) / M ** 3


def data_term(data, H_, K_, L_, ac_support, x0, M, rlambda, eps=1e-12):
    """d = adjoint(data, support) + rlambda x0 (sequential/autocorrelation.py:183-191)."""
    uvect_ADb = es_nufft.adjoint(np.ravel(data), H_, K_, L_, M, True, ac_support, eps).reshape(-1)
    return uvect_ADb + rlambda * x0


def cg(matvec, b, x0, maxiter, tol=1e-5, callback=None):
    """Conjugate gradients as cupyx/scipy ``cg(A, b, x0, tol, maxiter, callback)``
    (identity preconditioner); returns (x, info, iterations)."""
    x = np.array(x0, dtype=np.float64, copy=True)
    b = np.asarray(b, dtype=np.float64)
    b_norm = np.linalg.norm(b)
    if b_norm == 0:
        return b.copy(), 0, 0
    atol = tol * b_norm
    r = b - matvec(x)
    rho = 0.0
    p = None
    it = 0
    resid = np.linalg.norm(r)
    while it < maxiter:
        z = r
        rho1, rho = rho, float(r @ z)
        p = z.copy() if it == 0 else z + (rho / rho1) * p
        q = matvec(p)
        alpha = rho / float(p @ q)
        x = x + alpha * p
        r = r - alpha * q
        it += 1
        if callback is not None:
            callback(x)
        resid = np.linalg.norm(r)
        if resid <= atol:
            break
    info = it if (it == maxiter and not resid <= atol) else 0
    return x, info, it


def _reflect_index(n, radius):
    """Indices of scipy's mode 'reflect' (d c b a | a b c d | d c b a)."""
    i = np.arange(-radius, n + radius)
    period = 2 * n
    i = np.mod(i, period)
    return np.where(i >= n, period - 1 - i, i)


def gaussian_filter(a, sigma=0.5, truncate=4.0):
    """scipy.ndimage.gaussian_filter(a, sigma) in FP64 (separable, 'reflect')."""
    radius = int(truncate * float(sigma) + 0.5)
    x = np.arange(-radius, radius + 1, dtype=np.float64)
    w = np.exp(-0.5 * (x / sigma) ** 2)
    w /= w.sum()
    out = np.asarray(a, dtype=np.float64)
    for ax in range(out.ndim):
        n = out.shape[ax]
        pad = np.take(out, _reflect_index(n, radius), axis=ax)
        acc = np.zeros_like(out)
        for t in range(2 * radius + 1):
            acc += w[t] * np.take(pad, np.arange(t, t + n), axis=ax)
        out = acc
    return out


def solve_ac(slices_, H_, K_, L_, M, M_ups, oversampling, maxiter, ac_estimate=None,
             eps=1e-12, tol=1e-5, return_details=False):
    """Merge.solve_ac (sequential/autocorrelation.py:192-225) for given
    normalised positions H_, K_, L_ (= H * mult, :150-152) of the images'
    pixels; slices_ (N_images, ...) data."""
    N = int(np.prod(np.shape(slices_)))
    rlambda = 1 / N / 1000
    flambda = 1e3
    F_antisupport = antisupport(M, oversampling)
    if ac_estimate is None:
        ac_support = np.ones((M,) * 3)
        ac_estimate = np.zeros((M,) * 3)
    else:
        ac_smoothed = gaussian_filter(ac_estimate, 0.5)
        ac_support = (ac_smoothed > 1e-12).astype(np.float64)
        ac_estimate = ac_estimate * ac_support
    x0 = ac_estimate.reshape(-1)
    _, F_ugrid_conv_ = conv_kernel(H_, K_, L_, M, M_ups, eps)

    def W_matvec(uvect):
        uvect_ADA = core_problem_convolution(uvect, M, F_ugrid_conv_, M_ups, ac_support)
        uvect_FDF = fourier_reg(uvect, ac_support, F_antisupport, M)
        return uvect_ADA + rlambda * uvect + flambda * uvect_FDF

    d = data_term(np.asarray(slices_, np.float64).reshape(-1), H_, K_, L_, ac_support, x0, M,
                  rlambda, eps)
    ret, info, it = cg(W_matvec, d, x0, maxiter, tol)
    ac = ret.reshape((M,) * 3)
    if return_details:
        return ac, dict(info=info, iterations=it, support=ac_support, d=d,
                        F_ugrid_conv=F_ugrid_conv_, rlambda=rlambda, W=W_matvec)
    return ac
