"""Pins the oracle's quaternion -> rotation and q-point geometry."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import geometry
from spinifel_amd import quaternions, synth


def test_known_answer_skopi():
    """spinifel/tests/skopi_quaternion.py:11-32: [0, 0, 0, 1] <-> diag(-1, -1, 1)."""
    np.testing.assert_array_equal(geometry.quaternion2rot3d(np.array([0.0, 0, 0, 1])),
                                  np.diag([-1.0, -1.0, 1.0]))
    np.testing.assert_array_equal(geometry.quaternion2rot3d(np.array([1.0, 0, 0, 0])), np.eye(3))


def test_matches_scipy_active_rotation():
    """[w, x, y, z] convention equals scipy's from_quat([x, y, z, w]) (independent library)."""
    q = synth.random_quaternions(50, np.random.default_rng(0))
    R = geometry.quaternion2rot3d(q)
    S = Rotation.from_quat(q[:, [1, 2, 3, 0]]).as_matrix()
    np.testing.assert_allclose(R, S, atol=1e-14)


def test_invariants():
    q = synth.random_quaternions(100, np.random.default_rng(1))
    R = geometry.quaternion2rot3d(q)
    np.testing.assert_allclose(np.einsum("nij,nkj->nik", R, R), np.broadcast_to(np.eye(3), R.shape),
                               atol=1e-14)
    np.testing.assert_allclose(np.linalg.det(R), 1.0, atol=1e-14)
    np.testing.assert_allclose(geometry.quaternion2rot3d(-q), R, atol=1e-15)
    Rinv = geometry.ref_rotmat(q)
    np.testing.assert_allclose(Rinv, np.transpose(R, (0, 2, 1)), atol=1e-14)


def test_identity_positions_closed_form():
    ppr, pdr = synth.detector(2, 8, 16)
    E = pdr.max()
    H, K, L = geometry.gen_nonuniform_normalized_positions(np.array([[1.0, 0, 0, 0]]), ppr, E, 1)
    np.testing.assert_allclose(H, ppr[0].ravel() * np.pi / E, rtol=1e-15)
    np.testing.assert_allclose(K, ppr[1].ravel() * np.pi / E, rtol=1e-15)
    np.testing.assert_allclose(L, ppr[2].ravel() * np.pi / E, rtol=1e-15)
    # the corner pixel lands at |h| = 0.9 pi by construction of pdr
    r = np.sqrt(H ** 2 + K ** 2 + L ** 2)
    assert r.max() == pytest.approx(0.9 * np.pi, rel=1e-12)


def test_positions_rotate_with_inverse():
    q = synth.random_quaternions(3, np.random.default_rng(2))
    ppr, pdr = synth.detector(1, 4, 4)
    H, K, L = geometry.gen_nonuniform_positions(q, ppr)
    R = geometry.quaternion2rot3d(q)
    want = np.einsum("nji,jp->nip", R, ppr.reshape(3, -1))     # R^T q
    got = np.stack([H.reshape(3, -1), K.reshape(3, -1), L.reshape(3, -1)], 1)
    np.testing.assert_allclose(got, want, atol=1e-15)


def test_uniform_grid_is_unit_and_spread():
    q = quaternions.uniform_quaternions(2000)
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-14)
    assert (q[:, 0] >= 0).all()
    # near-uniform: largest gap (covering radius on SO(3)) is small
    rng = np.random.default_rng(3)
    probe = synth.random_quaternions(500, rng)
    d = np.abs(probe @ q.T).max(1)
    assert np.arccos(np.clip(d, -1, 1)).max() < 0.25
