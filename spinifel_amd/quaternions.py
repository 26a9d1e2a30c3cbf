"""Quaternion utilities ([w, x, y, z], the skopi convention).

``uniform_quaternions`` is the default reference-orientation grid when
``match`` gets ``ref_orientations=None``.  The reference uses
``skopi.get_uniform_quat(N, True)`` (spinifel/sequential/orientation_matching.py:217);
skopi is not available, so this is a different, deterministic
near-uniform grid: super-Fibonacci spirals (M. Alexa, "Super-Fibonacci
Spirals: Fast, Low-Discrepancy Sampling of SO(3)", CVPR 2022), folded to
the w >= 0 hemisphere.  Results that depend on the default grid are
therefore not comparable point-by-point with the reference's.
"""
import numpy as np

_PHI = np.sqrt(2.0)
_PSI = 1.533751168755204288118041


def uniform_quaternions(n):
    """n near-uniform unit quaternions (n, 4) float64, w >= 0."""
    n = int(n)
    s = np.arange(n, dtype=np.float64) + 0.5
    t = s / n
    r = np.sqrt(t)
    R = np.sqrt(1.0 - t)
    a = 2.0 * np.pi * s / _PHI
    b = 2.0 * np.pi * s / _PSI
    q = np.stack([r * np.sin(a), r * np.cos(a), R * np.sin(b), R * np.cos(b)], axis=1)
    # order as [w, x, y, z] with w the last spiral component, then fold
    q = q[:, [3, 0, 1, 2]]
    q[q[:, 0] < 0] *= -1.0
    return q


def random_quaternions(n, rng):
    """Uniform random unit quaternions (normalised 4-D Gaussians), w >= 0."""
    q = rng.standard_normal((int(n), 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 0] < 0] *= -1.0
    return q


def _rotvec(q):
    """Rotation vectors (axis * angle, |r| <= pi) of unit quaternions w >= 0."""
    w = np.clip(q[:, 0], -1.0, 1.0)
    v = q[:, 1:]
    s = np.linalg.norm(v, axis=1)
    ang = 2.0 * np.arctan2(s, w)
    axis = v / np.maximum(s, 1e-300)[:, None]
    return axis * ang[:, None]


def morton_order(q, bits=10):
    """Permutation sorting quaternions by the Morton code of their rotation vectors."""
    r = _rotvec(np.asarray(q, np.float64))
    u = np.clip(((r + np.pi) / (2.0 * np.pi) * (1 << bits)).astype(np.int64), 0, (1 << bits) - 1)
    code = np.zeros(len(u), dtype=np.int64)
    for b in range(bits):
        for d in range(3):
            code |= ((u[:, d] >> b) & 1) << (3 * b + d)
    return np.argsort(code, kind="stable")
