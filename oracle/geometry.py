"""ORACLE -- test infrastructure, not product code.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module.

FP64 restatement of the geometry spinifel applies before slicing:

* ``quaternion2rot3d``: skopi's quaternion -> rotation formula, quaternion
  ``[w, x, y, z]``; skopi is un-vendored, the formula is the in-tree copy in
  eval/geom.py:26-46 ("adopted from skopi").  Known answer
  ``q = [0, 0, 0, 1] -> diag(-1, -1, 1)`` from spinifel/tests/skopi_quaternion.py:11-32.
* ``ref_rotmat``: ``[inv(quaternion2rot3d(q)) for q in ref_orientations]``
  (spinifel/sequential/orientation_matching.py:220; same at
  spinifel/autocorrelation.py:294-295).
* ``gen_nonuniform_positions``: ``einsum("ijk,klmn->jilmn", rotmat, ppr)``
  (spinifel/autocorrelation.py:290-305, orientation_matching.py:233).
* ``gen_nonuniform_normalized_positions``: ``H.flatten() / E * pi / oversampling``
  (spinifel/autocorrelation.py:310-323, orientation_matching.py:234-236).

Parity status: pinned by the known answer above and by invariants
(orthonormal, det = +1, R(q) = R(-q), identity quaternion -> identity) in
tests/test_oracle_geometry.py, plus the closed form ``h = pi q / E`` for the
identity rotation.
"""
import numpy as np


def quaternion2rot3d(quat):
    """Rotation matrix of quaternion(s) ``[w, x, y, z]`` (eval/geom.py:26-46)."""
    q = np.asarray(quat, dtype=np.float64)
    single = q.ndim == 1
    q = np.atleast_2d(q)
    q01 = q[:, 0] * q[:, 1]
    q02 = q[:, 0] * q[:, 2]
    q03 = q[:, 0] * q[:, 3]
    q11 = q[:, 1] * q[:, 1]
    q12 = q[:, 1] * q[:, 2]
    q13 = q[:, 1] * q[:, 3]
    q22 = q[:, 2] * q[:, 2]
    q23 = q[:, 2] * q[:, 3]
    q33 = q[:, 3] * q[:, 3]
    r = np.zeros((q.shape[0], 3, 3))
    r[:, 0, 0] = 1.0 - 2.0 * (q22 + q33)
    r[:, 0, 1] = 2.0 * (q12 - q03)
    r[:, 0, 2] = 2.0 * (q13 + q02)
    r[:, 1, 0] = 2.0 * (q12 + q03)
    r[:, 1, 1] = 1.0 - 2.0 * (q11 + q33)
    r[:, 1, 2] = 2.0 * (q23 - q01)
    r[:, 2, 0] = 2.0 * (q13 - q02)
    r[:, 2, 1] = 2.0 * (q23 + q01)
    r[:, 2, 2] = 1.0 - 2.0 * (q11 + q22)
    return r[0] if single else r


def ref_rotmat(quats):
    """Inverse rotation per orientation, (N, 3, 3) FP64."""
    quats = np.asarray(quats, dtype=np.float64).reshape(-1, 4)
    if quats.shape[0] == 0:
        return np.zeros((0, 3, 3))
    return np.array([np.linalg.inv(quaternion2rot3d(q)) for q in quats])


def gen_nonuniform_positions(orientations, pixel_position_reciprocal):
    """H, K, L with shape (N_orientations,) + det_shape."""
    rot = ref_rotmat(orientations)
    ppr = np.asarray(pixel_position_reciprocal, dtype=np.float64)
    H, K, L = np.einsum("ijk,k...->ji...", rot, ppr)
    return H, K, L


def gen_nonuniform_normalized_positions(orientations, pixel_position_reciprocal,
                                        reciprocal_extent, oversampling):
    """Flattened H_, K_, L_ scaled by pi / (E * oversampling)."""
    H, K, L = gen_nonuniform_positions(orientations, pixel_position_reciprocal)
    H_ = H.flatten() / reciprocal_extent * np.pi / oversampling
    K_ = K.flatten() / reciprocal_extent * np.pi / oversampling
    L_ = L.flatten() / reciprocal_extent * np.pi / oversampling
    return H_, K_, L_
