"""Host-side pieces of the batching form: children CSR, Morton order and
batch sizes (no GPU)."""
import numpy as np

from spinifel_amd import hierarchy
from spinifel_amd.quaternions import uniform_quaternions


def test_children_csr_inverts_parent_lists():
    rng = np.random.default_rng(0)
    n_c, n_f, k = 17, 200, 4
    par = np.stack([rng.choice(n_c, k, replace=False) for _ in range(n_f)])
    off, ch = hierarchy.children_csr(par, n_c)
    assert off[0] == 0 and off[-1] == n_f * k and len(off) == n_c + 1
    for c in range(n_c):
        kids = ch[off[c]:off[c + 1]]
        assert np.all(np.diff(kids) > 0)                       # ascending, no repeats
        np.testing.assert_array_equal(kids, np.nonzero((par == c).any(1))[0])


def test_morton_order_is_a_local_permutation():
    q = uniform_quaternions(4000)
    o = hierarchy.morton_order(q)
    assert sorted(o.tolist()) == list(range(4000))
    qs = q[o]
    # consecutive nodes are much closer than random pairs
    d_seq = 1 - np.abs((qs[1:] * qs[:-1]).sum(1))
    rng = np.random.default_rng(1)
    a, b = rng.integers(0, 4000, (2, 3999))
    d_rand = 1 - np.abs((q[a] * q[b]).sum(1))
    assert np.median(d_seq) < 0.1 * np.median(d_rand)


def test_rotvec_norm_is_angle():
    q = uniform_quaternions(500)
    r = hierarchy._rotvec(q)
    ang = 2 * np.arccos(np.clip(q[:, 0], -1, 1))
    np.testing.assert_allclose(np.linalg.norm(r, axis=1), ang, atol=1e-12)


def test_log_uniform_batches_range():
    rng = np.random.default_rng(2)
    s = hierarchy.log_uniform_batches(10000, 16, 4096, rng)
    assert s.min() >= 16 and s.max() < 4096
    # log-uniform: median near sqrt(16 * 4096) = 256
    assert 200 < np.median(s) < 330


def test_twin_is_half_turn_about_beam_and_involutive():
    q = uniform_quaternions(300)
    t = hierarchy.twin_quaternions(q)
    tt = hierarchy.twin_quaternions(t)
    np.testing.assert_allclose(np.abs((tt * q).sum(1)), 1.0, atol=1e-12)   # twin of twin = q
    # R(twin) = Rz(pi) R(q): relative rotation is a half turn about z
    rel = hierarchy.quat_mul(t, q * np.array([1, -1, -1, -1]))
    np.testing.assert_allclose(np.abs(rel[:, 3]), 1.0, atol=1e-12)


def test_union_csr_rows():
    off = np.array([0, 2, 3, 6], np.int32)
    ch = np.array([1, 4, 2, 0, 4, 9], np.int32)
    o, c = hierarchy.union_csr(off, ch, [2, 0, 1])
    rows = [c[o[i]:o[i + 1]].tolist() for i in range(3)]
    assert rows == [[0, 1, 4, 9], [1, 2, 4], [0, 2, 4, 9]]
