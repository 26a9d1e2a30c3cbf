"""Pins the oracle's type-2 NUFFT (FINUFFT method restated) with closed forms
and brute force, independently of its own formula."""
import numpy as np
import pytest

from helpers import rel_l2
from oracle import es_nufft


def _rand_model(M, seed, sym=True):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((M, M, M))
    if sym and M % 2 == 1:
        f = f + f[::-1, ::-1, ::-1]
    return f


def _pts(n, seed, lim=np.pi):
    return np.random.default_rng(seed).uniform(-lim, lim, (3, n))


def test_es_params_follow_finufft_rule():
    assert es_nufft.es_params(1e-6) == (7, pytest.approx(2.30 * 7))
    assert es_nufft.es_params(1e-12)[0] == 13
    assert es_nufft.es_params(1e-3)[0] == 4
    assert es_nufft.es_params(1e-3)[1] == pytest.approx(2.38 * 4)
    assert es_nufft.next235even(82) == 90
    assert es_nufft.next235even(162) == 162
    assert es_nufft.next235even(242) == 250


@pytest.mark.parametrize("eps,tol", [(1e-6, 2e-6), (1e-12, 1e-11)])
@pytest.mark.parametrize("M", [15, 16])
def test_nufft_matches_direct_sum(eps, tol, M):
    f = _rand_model(M, M)
    x = _pts(300, 1)
    c = es_nufft.nufft3d2(x[0], x[1], x[2], f, -1, eps)
    ref = es_nufft.nudft3d2_numpy(x[0], x[1], x[2], f, -1)
    assert np.linalg.norm(c - ref) / np.linalg.norm(ref) < tol


def test_direct_sum_c_equals_numpy():
    f = _rand_model(11, 3, sym=False)
    x = _pts(50, 4)
    a = es_nufft.nudft3d2(x[0], x[1], x[2], f, -1)
    b = es_nufft.nudft3d2_numpy(x[0], x[1], x[2], f, -1)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)
    # isign = +1 is the complex conjugate for a real model
    c = es_nufft.nudft3d2(x[0], x[1], x[2], f, +1)
    np.testing.assert_allclose(c, np.conj(a), rtol=1e-12, atol=1e-12)


def test_interp_c_equals_numpy():
    f = _rand_model(13, 5)
    b, n, w, beta = es_nufft.fine_grid(f, -1, 1e-6)
    x = _pts(200, 6, 2.9 * np.pi)
    np.testing.assert_allclose(es_nufft.interp(b, x[0], x[1], x[2], w, beta),
                               es_nufft.interp_numpy(b, x[0], x[1], x[2], w, beta),
                               rtol=1e-12, atol=1e-12)


def test_delta_at_centre_is_constant():
    M = 17
    f = np.zeros((M, M, M))
    f[M // 2, M // 2, M // 2] = 3.0
    x = _pts(200, 7)
    c = es_nufft.nufft3d2(x[0], x[1], x[2], f, -1, 1e-9)
    np.testing.assert_allclose(c.real, 3.0, rtol=1e-8)
    np.testing.assert_allclose(c.imag, 0.0, atol=1e-8)


def test_symmetric_pair_is_cosine():
    M = 21
    k0 = np.array([3, -5, 7])
    f = np.zeros((M, M, M))
    f[tuple(M // 2 + k0)] = 1.0
    f[tuple(M // 2 - k0)] = 1.0
    x = _pts(200, 8)
    c = es_nufft.nufft3d2(x[0], x[1], x[2], f, -1, 1e-9)
    want = 2.0 * np.cos(k0 @ x)
    np.testing.assert_allclose(c.real, want, atol=1e-8 * 2)
    np.testing.assert_allclose(c.imag, 0.0, atol=1e-8)


def test_lattice_points_equal_fft():
    """x = 2 pi m / M: sum_k f_k e^{-i k.x} = e^{i h 2pi m / M} fftn(f)[m], h = M // 2."""
    M = 12
    f = _rand_model(M, 9, sym=False)
    F = np.fft.fftn(f)
    m = np.random.default_rng(10).integers(0, M, (3, 40))
    x = 2 * np.pi * m / M
    x = np.where(x >= np.pi, x - 2 * np.pi, x)
    c = es_nufft.nufft3d2(x[0], x[1], x[2], f, -1, 1e-12)
    h = M // 2
    want = np.exp(1j * 2 * np.pi * h * m.sum(0) / M) * F[m[0], m[1], m[2]]
    np.testing.assert_allclose(c, want, rtol=1e-9, atol=1e-9 * np.abs(F).max())


def test_axis_pairing_permutation():
    """Axis 0 of the model pairs with x: transposing the model equals permuting the points."""
    f = _rand_model(13, 11, sym=False)
    x = _pts(100, 12)
    a = es_nufft.nufft3d2(x[0], x[1], x[2], f, -1, 1e-9)
    b = es_nufft.nufft3d2(x[2], x[0], x[1], np.transpose(f, (2, 0, 1)), -1, 1e-9)
    np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-8 * np.abs(a).max())


def test_centrosymmetric_model_gives_real_transform():
    f = _rand_model(15, 13, sym=True)
    x = _pts(200, 14)
    c = es_nufft.nufft3d2(x[0], x[1], x[2], f, -1, 1e-9)
    assert np.abs(c.imag).max() < 1e-7 * np.abs(c.real).max()


def test_domain_assert():
    f = _rand_model(9, 1)
    with pytest.raises(AssertionError):
        es_nufft.nufft3d2([3.2 * np.pi], [0.0], [0.0], f)


def test_periodic_wrap():
    f = _rand_model(11, 2)
    x = _pts(100, 3)
    a = es_nufft.nufft3d2(x[0], x[1], x[2], f, -1, 1e-9)
    b = es_nufft.nufft3d2(x[0] + 2 * np.pi, x[1], x[2] - 2 * np.pi, f, -1, 1e-9)
    np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-8 * np.abs(a).max())
