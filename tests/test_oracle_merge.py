"""Pins the oracle's merge (solve_ac) pieces against what they stand for:
the convolution trick against brute-force A^H A, the Fourier regulariser by
projection properties and a closed form, CG and the Gaussian filter against
scipy (importable here), and the solve against its own linear system."""
import numpy as np
import pytest

from helpers import rel_l2
from oracle import es_nufft, merge


def _pts(n, seed, lim=np.pi / 2):
    return np.random.default_rng(seed).uniform(-lim, lim, (3, n))


@pytest.mark.parametrize("M", [9, 10])
def test_convolution_equals_brute_force_AHA(M):
    """core_problem_convolution(u) * M_ups^3 = s . Re A^H A (s . u) / M^3 with A
    the type-2 direct sum at the points and A^H the type-1 direct sum."""
    M_ups = 2 * M
    x = _pts(300, M)
    rng = np.random.default_rng(M + 1)
    u = rng.standard_normal(M ** 3)
    s = (rng.uniform(size=(M,) * 3) > 0.3).astype(np.float64)
    _, F = merge.conv_kernel(*x, M, M_ups, 1e-12)
    conv = merge.core_problem_convolution(u, M, F, M_ups, s)
    fwd = es_nufft.nudft3d2(*x, u.reshape((M,) * 3) * s, -1)
    aha = (es_nufft.nudft3d1(*x, fwd, M, +1) * s).real / M ** 3
    assert rel_l2(conv * M_ups ** 3, aha) < 1e-10


def test_fourier_reg_projection_properties():
    M = 11
    rng = np.random.default_rng(3)
    u = rng.standard_normal(M ** 3)
    ones = np.ones((M,) * 3)
    A = merge.antisupport(M, 2)
    assert np.all(A == A[::-1, ::-1, ::-1]) and A.any() and not A.all()
    p1 = merge.fourier_reg(u, ones, A, M)
    p2 = merge.fourier_reg(p1, ones, A, M)
    assert rel_l2(p2, p1) < 1e-12
    np.testing.assert_allclose(merge.fourier_reg(u, ones, np.ones_like(A), M), u, atol=1e-12)
    np.testing.assert_allclose(merge.fourier_reg(u, ones, np.zeros_like(A), M), 0, atol=1e-12)


def test_fourier_reg_delta_closed_form():
    """A delta at the centre index M//2 gives (1/M^3) sum_m mask(m) cos(2 pi m.(j - c) / M)
    with mask(m) the antisupport at frequency m (ifftshift order)."""
    M = 9
    A = merge.antisupport(M, 2)
    u = np.zeros((M,) * 3)
    c = M // 2
    u[c, c, c] = 1.0
    out = merge.fourier_reg(u.reshape(-1), np.ones((M,) * 3), A, M).reshape((M,) * 3)
    As = np.fft.ifftshift(A).astype(np.float64)
    m = np.arange(M)
    j = np.arange(M) - c
    E = np.exp(2j * np.pi * np.outer(m, j) / M)       # (m, j)
    ref = np.einsum("abc,ai,bj,ck->ijk", As, E, E, E).real / M ** 3
    np.testing.assert_allclose(out, ref, atol=1e-12)


def test_cg_matches_scipy_and_solve():
    from scipy.sparse.linalg import cg as scipy_cg
    rng = np.random.default_rng(5)
    n = 60
    Q = rng.standard_normal((n, n))
    A = Q @ Q.T + n * np.eye(n)
    b = rng.standard_normal(n)
    x0 = rng.standard_normal(n)
    calls = []
    x, info, it = merge.cg(lambda v: A @ v, b, x0, 500, 1e-10, callback=lambda xk: calls.append(1))
    assert info == 0 and len(calls) == it
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)
    try:
        xs, _ = scipy_cg(A, b, x0=x0, maxiter=500, rtol=1e-10)
    except TypeError:
        xs, _ = scipy_cg(A, b, x0=x0, maxiter=500, tol=1e-10)
    np.testing.assert_allclose(x, xs, rtol=1e-7, atol=1e-9)
    # the stopping rule: ||r|| <= tol ||b|| reached at the reported iteration
    assert np.linalg.norm(b - A @ x) <= 1e-10 * np.linalg.norm(b) * 1.0001
    # maxiter caps the updates and reports info = maxiter when not converged
    _, info3, it3 = merge.cg(lambda v: A @ v, b, x0, 3, 1e-14)
    assert it3 == 3 and info3 == 3


def test_gaussian_filter_matches_scipy():
    from scipy.ndimage import gaussian_filter
    a = np.random.default_rng(7).standard_normal((9, 10, 11))
    np.testing.assert_allclose(merge.gaussian_filter(a, 0.5), gaussian_filter(a, 0.5),
                               rtol=1e-12, atol=1e-14)


def test_solve_ac_solves_its_system():
    """With enough iterations the returned ac satisfies W ac = d to CG's
    tolerance; with an estimate, its support is where the smoothed estimate
    exceeds 1e-12 and ac vanishes outside it."""
    M, M_ups = 9, 18
    x = _pts(400, 11)
    data = np.random.default_rng(12).uniform(0, 1, (4, 100))
    ac, det = merge.solve_ac(data, *x, M, M_ups, 2, 200, return_details=True)
    r = det["W"](ac.reshape(-1)) - det["d"]
    assert det["info"] == 0
    assert np.linalg.norm(r) <= 1e-5 * np.linalg.norm(det["d"]) * 1.0001
    est = np.zeros((M,) * 3)
    est[4, 4, 4] = abs(ac[4, 4, 4]) + 1.0      # smoothed support: the 5^3 cube around it
    ac2, det2 = merge.solve_ac(data, *x, M, M_ups, 2, 20, ac_estimate=est, return_details=True)
    sup = det2["support"]
    assert sup.sum() == 125
    assert np.all(ac2[sup == 0] == 0)
