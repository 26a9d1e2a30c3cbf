"""Pins the oracle's type-1 (adjoint) NUFFT with brute force, the adjoint
identity against the already-pinned type-2 brute force, and closed forms --
none of which retypes the spreading formula."""
import numpy as np
import pytest

from helpers import rel_l2
from oracle import es_nufft


def _pts(n, seed, lim=np.pi):
    return np.random.default_rng(seed).uniform(-lim, lim, (3, n))


def _vals(n, seed, complex_=True):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n)
    return v + 1j * rng.standard_normal(n) if complex_ else v


def test_direct_type1_c_equals_numpy():
    x = _pts(60, 1)
    c = _vals(60, 2)
    for M in (7, 8):
        for s in (1, -1):
            np.testing.assert_allclose(es_nufft.nudft3d1(*x, c, M, s),
                                       es_nufft.nudft3d1_numpy(*x, c, M, s),
                                       rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("M", [9, 10])
def test_adjoint_identity_with_type2(M):
    """<A f, c> = <f, A^H c> with A the type-2 sum (isign -1, pinned by its own
    tests) and A^H the type-1 sum (isign +1): a sign, index or axis slip in the
    type-1 sum breaks it."""
    rng = np.random.default_rng(M)
    x = _pts(200, 3)
    c = _vals(200, 4)
    f = rng.standard_normal((M, M, M))
    Af = es_nufft.nudft3d2(*x, f, -1)
    AHc = es_nufft.nudft3d1(*x, c, M, +1)
    lhs = np.vdot(c, Af)
    rhs = np.sum(f * np.conj(AHc))
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)
    # and the ES type 1 satisfies it to its tolerance
    AHc_es = es_nufft.nufft3d1(*x, c, M, +1, 1e-12)
    assert abs(np.sum(f * np.conj(AHc_es)) - lhs) <= 1e-10 * abs(lhs)


def test_point_at_origin_is_constant():
    M = 8
    x = np.zeros((3, 1))
    out = es_nufft.nufft3d1(*x, np.array([2.5 - 1j]), M, 1, 1e-12)
    np.testing.assert_allclose(out, np.full((M,) * 3, 2.5 - 1j), rtol=1e-11, atol=1e-11)


def test_lattice_points_equal_inverse_fft():
    """Points on 2 pi m / M with values g_m: f_k = sum_m g_m exp(+2 pi i k m / M),
    i.e. M^3 ifftn(g) read at k mod M."""
    M = 8
    rng = np.random.default_rng(7)
    g = rng.standard_normal((M, M, M)) + 1j * rng.standard_normal((M, M, M))
    m = np.arange(M)
    m1, m2, m3 = np.meshgrid(m, m, m, indexing="ij")
    xs = [2 * np.pi * mm.ravel() / M for mm in (m1, m2, m3)]
    xs = [np.where(v >= np.pi, v - 2 * np.pi, v) for v in xs]
    out = es_nufft.nufft3d1(*xs, g.ravel(), M, 1, 1e-12)
    k = np.arange(M) - M // 2
    ref = (np.fft.ifftn(g) * M ** 3)[np.ix_(k % M, k % M, k % M)]
    assert rel_l2(out.real, ref.real) < 1e-11 and rel_l2(out.imag, ref.imag) < 1e-11


@pytest.mark.parametrize("eps,tol", [(1e-6, 2e-6), (1e-12, 1e-11)])
@pytest.mark.parametrize("M", [15, 16])
def test_es_type1_matches_direct_sum(eps, tol, M):
    x = _pts(400, M, 2.9 * np.pi)
    c = _vals(400, M + 1)
    ref = es_nufft.nudft3d1_numpy(*x, c, M, 1)
    out = es_nufft.nufft3d1(*x, c, M, 1, eps)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < tol


def test_adjoint_dropin_semantics():
    """spinifel's adjoint: support applied, real part kept, divided by M^3."""
    M = 9
    x = _pts(100, 11)
    c = _vals(100, 12, complex_=False)
    sup = (np.random.default_rng(13).uniform(size=(M,) * 3) > 0.3).astype(np.float64)
    full = es_nufft.nudft3d1(*x, c, M, 1)
    out = es_nufft.adjoint(c, *x, M, True, sup)
    assert out.dtype == np.float64
    np.testing.assert_allclose(out, (full * sup).real / M ** 3, rtol=0, atol=1e-10)
    with pytest.raises(AssertionError):
        es_nufft.adjoint(c, x[0] * 4, x[1], x[2], M)
