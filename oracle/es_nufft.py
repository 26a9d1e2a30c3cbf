"""ORACLE -- test infrastructure, not product code.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only as the checker / the timed CPU baseline.

FP64 restatement of the type-2 3D NUFFT that spinifel calls as
``nufft_3d_t2(H_, K_, L_, ugrid, -1, 1e-12, N)`` (spinifel/autocorrelation.py:59,
alias spinifel/extern/__init__.py:79-83 -> finufftpy 1.1.2 ``nufft3d2`` in
spinifel/extern/finufft_ext.py:70-94).  finufft is un-vendored in the
reference, so this follows its published exponential-of-semicircle method:

1. width ``w = ceil(-log10(eps/10))`` (upsampling sigma = 2), ES parameter
   ``beta = 2.30 w`` (2.20, 2.26, 2.38 for w = 2, 3, 4);
2. fine grid ``n = next235even(max(sigma*M, 2w))``;
3. deconvolve: ``f_k / (phihat(k1) phihat(k2) phihat(k3))`` with
   ``phihat(k) = int phi(x) exp(-2 pi i k x / n) dx`` by Gauss-Legendre
   quadrature; zero-pad to n^3 with mode k at index ``k mod n``;
4. ``b = FFT_n(.)`` with the sign of ``isign``;
5. interpolate ``c_j = sum_l b_l prod_d phi(X_d - l_d)`` (C, ``esnufft.c``).

Mode order is finufft's default CMCL: model index ``i`` along each axis is
frequency ``k = i - M//2``.  Axis 0 pairs with x (= H).

The type-1 (adjoint) half, ``nufft3d1`` / ``adjoint``, spreads with the same
kernel on the same fine grid (C, ``oracle_es_spread3d``), FFTs with the sign
of ``isign`` and deconvolves; ``nudft3d1`` is its brute-force ground truth.

Parity status: pinned by closed forms and brute force (tests/test_oracle_nufft.py):
the FP64 direct NUDFT below, a delta at the centre (constant), a symmetric
pair (2 cos), points on the lattice 2 pi m / M (equals numpy's FFT), axis
permutation, and realness for centrosymmetric models.  Type 1
(tests/test_oracle_adjoint.py): the brute-force sum against numpy, against
the type-2 brute force through the adjoint identity <A f, c> = <f, A^H c>, a
point at the origin (constant), lattice points (numpy's inverse FFT), and the
ES type 1 against the brute force.  finufft's own output
bits are not reproducible here (library absent); only tolerance-level
agreement is meaningful for an approximate transform anyway.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    """Compile the C part of the oracle (``make -C oracle``)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            raise RuntimeError(
                f"oracle C library missing: {_LIBPATH} (run `make -C oracle`)")
        lib = ctypes.CDLL(_LIBPATH)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.oracle_es_interp3d.argtypes = [dp, ctypes.c_int64, dp, dp, dp,
                                           ctypes.c_int64, ctypes.c_int,
                                           ctypes.c_double, dp]
        lib.oracle_es_interp3d.restype = ctypes.c_int
        lib.oracle_nudft3d2_real.argtypes = [dp, ctypes.c_int64, dp, dp, dp,
                                             ctypes.c_int64, ctypes.c_int, dp]
        lib.oracle_nudft3d2_real.restype = ctypes.c_int
        lib.oracle_es_spread3d.argtypes = [dp, dp, dp, dp, ctypes.c_int64, ctypes.c_int64,
                                           ctypes.c_int, ctypes.c_double, dp]
        lib.oracle_es_spread3d.restype = ctypes.c_int
        lib.oracle_nudft3d1.argtypes = [dp, dp, dp, dp, ctypes.c_int64, ctypes.c_int64,
                                        ctypes.c_int, dp]
        lib.oracle_nudft3d1.restype = ctypes.c_int
        _lib = lib
    return _lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def es_params(eps, upsampfac=2.0):
    """Kernel width and ES beta for a requested tolerance (finufft rule)."""
    if upsampfac != 2.0:
        raise ValueError("oracle implements sigma = 2 only")
    eps = max(float(eps), 1e-15)
    w = int(np.ceil(-np.log10(eps / 10.0)))
    w = max(2, min(16, w))
    beta_over_w = {2: 2.20, 3: 2.26, 4: 2.38}.get(w, 2.30)
    return w, beta_over_w * w


def next235even(n):
    """Smallest even integer >= n whose prime factors are 2, 3, 5."""
    n = int(n)
    if n <= 2:
        return 2
    if n % 2:
        n += 1
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 2


def fine_size(M, w, upsampfac=2.0):
    return next235even(max(int(np.ceil(upsampfac * M)), 2 * w))


def es_kernel(z, w, beta):
    """phi(z) = exp(beta (sqrt(1 - (2z/w)^2) - 1)) on |z| < w/2, else 0."""
    z = np.asarray(z, dtype=np.float64)
    t = 2.0 * z / w
    v = 1.0 - t * t
    out = np.zeros_like(z)
    m = v > 0
    out[m] = np.exp(beta * (np.sqrt(v[m]) - 1.0))
    return out


def es_fourier(k, n, w, beta, nquad=None):
    """phihat(k) = int_{-w/2}^{w/2} phi(x) cos(2 pi k x / n) dx (Gauss-Legendre)."""
    if nquad is None:
        nquad = 2 * (2 + 3 * int(np.ceil(w / 2.0))) + 32
    zq, wq = np.polynomial.legendre.leggauss(nquad)
    x = zq * (w / 2.0)
    wq = wq * (w / 2.0)
    ph = es_kernel(x, w, beta)
    k = np.asarray(k, dtype=np.float64)
    return (np.cos(2.0 * np.pi * np.outer(k, x) / n) * (ph * wq)[None, :]).sum(1)


def fine_grid(f, isign=-1, eps=1e-12, upsampfac=2.0):
    """Steps 1-4: complex fine grid b (n, n, n) plus (n, w, beta)."""
    f = np.asarray(f)
    M1, M2, M3 = f.shape
    if not (M1 == M2 == M3):
        raise ValueError("oracle handles cubic models only")
    M = M1
    w, beta = es_params(eps, upsampfac)
    n = fine_size(M, w, upsampfac)
    k = np.arange(M) - M // 2
    ph = es_fourier(k, n, w, beta)
    g = f.astype(np.complex128) / (ph[:, None, None] * ph[None, :, None] * ph[None, None, :])
    G = np.zeros((n, n, n), dtype=np.complex128)
    idx = k % n
    G[np.ix_(idx, idx, idx)] = g
    if isign < 0:
        b = np.fft.fftn(G)
    else:
        b = np.fft.ifftn(G) * (n ** 3)
    return np.ascontiguousarray(b), n, w, beta


def interp(b, x, y, z, w, beta):
    """Step 5 in C/OpenMP (complex128 output)."""
    lib = _load()
    b = np.ascontiguousarray(b, dtype=np.complex128)
    n = b.shape[0]
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    y = np.ascontiguousarray(y, dtype=np.float64).ravel()
    z = np.ascontiguousarray(z, dtype=np.float64).ravel()
    out = np.zeros(x.shape[0], dtype=np.complex128)
    bv = b.view(np.float64)
    ov = out.view(np.float64)
    rc = lib.oracle_es_interp3d(_dptr(bv), n, _dptr(x), _dptr(y), _dptr(z),
                                x.shape[0], int(w), float(beta), _dptr(ov))
    if rc:
        raise ValueError("oracle_es_interp3d: bad arguments")
    return out


def interp_numpy(b, x, y, z, w, beta):
    """Pure-numpy step 5 (small inputs; cross-checks the C loop)."""
    n = b.shape[0]
    X = np.stack([np.asarray(v, np.float64).ravel() for v in (x, y, z)], 1)
    X = X / (2 * np.pi)
    X = (X - np.floor(X)) * n
    i0 = np.ceil(X - w / 2.0).astype(np.int64)
    out = np.zeros(X.shape[0], dtype=np.complex128)
    for a in range(w):
        wa = es_kernel(i0[:, 0] + a - X[:, 0], w, beta)
        for bb in range(w):
            wb = es_kernel(i0[:, 1] + bb - X[:, 1], w, beta)
            for c in range(w):
                wc = es_kernel(i0[:, 2] + c - X[:, 2], w, beta)
                out += wa * wb * wc * b[(i0[:, 0] + a) % n, (i0[:, 1] + bb) % n,
                                        (i0[:, 2] + c) % n]
    return out


def nufft3d2(x, y, z, f, isign=-1, eps=1e-12, upsampfac=2.0):
    """c_j = sum_k f_k exp(isign i k.(x_j, y_j, z_j)), FP64 complex.

    Restates ``finufftpy.nufft3d2(x, y, z, c, isign, eps, f)``
    (spinifel/extern/finufft_ext.py:88).  Raises like
    spinifel/autocorrelation.py:46 when a coordinate leaves [-3 pi, 3 pi).
    """
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    z = np.asarray(z, np.float64)
    if x.size and max(np.abs(x).max(), np.abs(y).max(), np.abs(z).max()) >= 3 * np.pi:
        raise AssertionError("NU points outside [-3pi, 3pi)")
    b, n, w, beta = fine_grid(f, isign, eps, upsampfac)
    return interp(b, x, y, z, w, beta)


def nudft3d2(x, y, z, f, isign=-1):
    """FP64 direct non-uniform DFT (brute force; the ground truth), real f."""
    lib = _load()
    f = np.ascontiguousarray(f, dtype=np.float64)
    M = f.shape[0]
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    y = np.ascontiguousarray(y, dtype=np.float64).ravel()
    z = np.ascontiguousarray(z, dtype=np.float64).ravel()
    out = np.zeros(x.shape[0], dtype=np.complex128)
    rc = lib.oracle_nudft3d2_real(_dptr(f), M, _dptr(x), _dptr(y), _dptr(z),
                                  x.shape[0], int(isign), _dptr(out.view(np.float64)))
    if rc:
        raise ValueError("oracle_nudft3d2_real: bad arguments")
    return out


def nudft3d2_numpy(x, y, z, f, isign=-1):
    """Pure-numpy direct NUDFT (tiny inputs; cross-checks the C loop)."""
    f = np.asarray(f, np.float64)
    M = f.shape[0]
    k = np.arange(M) - M // 2
    e1 = np.exp(1j * isign * np.outer(np.ravel(x), k))
    e2 = np.exp(1j * isign * np.outer(np.ravel(y), k))
    e3 = np.exp(1j * isign * np.outer(np.ravel(z), k))
    return np.einsum("abc,ja,jb,jc->j", f, e1, e2, e3)


# ------------------------------------------------------------ type 1 (adjoint)

def _pts(x, y, z, c):
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    y = np.ascontiguousarray(y, dtype=np.float64).ravel()
    z = np.ascontiguousarray(z, dtype=np.float64).ravel()
    c = np.ascontiguousarray(np.asarray(c).ravel(), dtype=np.complex128)
    if not (x.size == y.size == z.size == c.size):
        raise ValueError("x, y, z, c differ in length")
    return x, y, z, c


def spread(x, y, z, c, n, w, beta):
    """b (n, n, n) complex: sum_j c_j prod_d phi(l_d - X_jd) on the periodic grid (C)."""
    lib = _load()
    x, y, z, c = _pts(x, y, z, c)
    grid = np.zeros((n, n, n), dtype=np.complex128)
    rc = lib.oracle_es_spread3d(_dptr(x), _dptr(y), _dptr(z), _dptr(c.view(np.float64)),
                                x.size, int(n), int(w), float(beta),
                                _dptr(grid.view(np.float64)))
    if rc:
        raise ValueError("oracle_es_spread3d: bad arguments")
    return grid


def nufft3d1(x, y, z, c, M, isign=1, eps=1e-12, upsampfac=2.0):
    """f_k = sum_j c_j exp(isign i k.(x_j, y_j, z_j)), k_d in [-M/2, (M-1)/2],
    complex (M, M, M), axis 0 with x.

    Restates ``nufft_3d_t1(H_, K_, L_, nuvect, +1, eps, M, M, M)``
    (spinifel/autocorrelation.py:119; finufftpy ``nufft3d1`` at
    spinifel/extern/finufft_ext.py:59): spread with the ES kernel on the fine
    grid, FFT with the sign of ``isign``, keep modes k mod n, divide by
    phihat(k1) phihat(k2) phihat(k3).
    """
    x, y, z, c = _pts(x, y, z, c)
    if x.size and max(np.abs(x).max(), np.abs(y).max(), np.abs(z).max()) >= 3 * np.pi:
        raise AssertionError("NU points outside [-3pi, 3pi)")
    w, beta = es_params(eps, upsampfac)
    n = fine_size(M, w, upsampfac)
    b = spread(x, y, z, c, n, w, beta)
    B = np.fft.ifftn(b) * n ** 3 if isign > 0 else np.fft.fftn(b)
    k = np.arange(M) - M // 2
    idx = k % n
    ph = es_fourier(k, n, w, beta)
    f = B[np.ix_(idx, idx, idx)]
    return f / (ph[:, None, None] * ph[None, :, None] * ph[None, None, :])


def nudft3d1(x, y, z, c, M, isign=1):
    """FP64 direct type-1 sum (brute force; the ground truth), complex (M, M, M)."""
    lib = _load()
    x, y, z, c = _pts(x, y, z, c)
    out = np.zeros((M, M, M), dtype=np.complex128)
    rc = lib.oracle_nudft3d1(_dptr(x), _dptr(y), _dptr(z), _dptr(c.view(np.float64)), x.size,
                             int(M), int(isign), _dptr(out.view(np.float64)))
    if rc:
        raise ValueError("oracle_nudft3d1: bad arguments")
    return out


def nudft3d1_numpy(x, y, z, c, M, isign=1):
    """Pure-numpy direct type-1 sum (tiny inputs; cross-checks the C loop)."""
    k = np.arange(M) - M // 2
    e1 = np.exp(1j * isign * np.outer(np.ravel(x), k))
    e2 = np.exp(1j * isign * np.outer(np.ravel(y), k))
    e3 = np.exp(1j * isign * np.outer(np.ravel(z), k))
    return np.einsum("j,ja,jb,jc->abc", np.ravel(c).astype(np.complex128), e1, e2, e3)


def adjoint(nuvect, H_, K_, L_, M, use_recip_sym=True, support=None, eps=1e-12):
    """spinifel's ``adjoint`` (spinifel/autocorrelation.py:97-128): type 1 with
    isign +1, times support, real part when use_recip_sym, divided by M^3."""
    ugrid = nufft3d1(H_, K_, L_, nuvect, M, 1, eps)
    if support is not None:
        ugrid = ugrid * support
    if use_recip_sym:
        ugrid = ugrid.real
    return ugrid / M ** 3
