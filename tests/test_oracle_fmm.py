"""Pins the oracle's Chebyshev interpolation, octree and black-box FMM
(fmm/chebyshev.rs, fmm/tree.rs, fmm/bbfmm.rs) with the reference's own unit
tests and the brute-force O(N^2) sum."""
import numpy as np
import pytest

from oracle import oracle as O


def test_lcg_point_set_is_the_reference_generator():
    # crates/pex/tests/fmm.rs:21-29, recomputed with Python integers
    pts = O.lcg_points(5, 12345)
    s = 12345
    vals = []
    for _ in range(15):
        s = (s * 6364136223846793005 + 1442695040888963407) % 2**64
        vals.append((s >> 33) / 2**31 - 1.0)
    assert np.array_equal(pts.ravel(), np.array(vals))


def test_interpolation_exact_for_low_degree():
    # chebyshev.rs:112-124
    n = 6
    nd = O.cheb_nodes(n)
    f = lambda x: 3 - 2 * x + x * x - 0.5 * x ** 3
    for xt in (-0.9, -0.3, 0.0, 0.42, 0.777):
        w = O.interp_weights(xt, n)
        assert abs(w @ f(nd) - f(xt)) < 1e-11


def test_interpolation_converges():
    # chebyshev.rs:126-141
    def err(n):
        nd = O.cheb_nodes(n)
        xs = np.linspace(-1, 1, 21)
        return max(abs(O.interp_weights(x, n) @ np.exp(nd) - np.exp(x)) for x in xs)

    assert err(8) < err(4) and err(8) < 1e-6


def test_tensor_weights_reproduce_trilinear():
    # chebyshev.rs:143-153 (3-D weights = outer product of 1-D ones)
    n = 5
    nd = O.cheb_nodes(n)
    f = lambda x, y, z: 1 + x - 2 * y + z + x * y * z
    pt = (0.3, -0.6, 0.1)
    w = [O.interp_weights(c, n) for c in pt]
    X, Y, Z = np.meshgrid(nd, nd, nd, indexing="ij")
    approx = np.einsum("a,b,c,abc->", w[0], w[1], w[2], f(X, Y, Z))
    assert abs(approx - f(*pt)) < 1e-10


def test_nodes_are_chebyshev_roots():
    for n in (3, 4, 5, 6):
        nd = O.cheb_nodes(n)
        # T_n(x_m) = cos(n arccos x_m) = 0
        assert np.allclose(np.cos(n * np.arccos(nd)), 0.0, atol=1e-14)


def grid_points(n):
    g = np.arange(n, dtype=float)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


def test_tree_covers_points_and_lists_are_separated():
    # tree.rs:244-269
    st = O.octree_stats(grid_points(4), 3)
    assert st["leaf_points"] == 64
    st = O.octree_stats(grid_points(8), 3)
    assert st["leaf_points"] == 512
    assert st["min_sep"] > 1.5


def test_dense_p2p_matches_numpy():
    pts = O.lcg_points(200, 7)
    q = np.linspace(-1, 1, 200)
    d = np.linalg.norm(pts[:, None, :] - pts[None, :, :], axis=2)
    np.fill_diagonal(d, np.inf)
    ref = (q[None, :] / d).sum(1)
    scale = (np.abs(q)[None, :] / d).sum(1)
    assert np.all(np.abs(O.dense_p2p(pts, q) - ref) <= 1e-13 * scale)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_fmm_matvec_matches_dense():
    # crates/pex/tests/fmm.rs:37-47
    pts = O.lcg_points(900, 12345)
    q = np.array([(i % 13) - 6.0 for i in range(900)])
    e = rel(O.fmm_matvec(pts, q, 5, 3), O.dense_p2p(pts, q))
    assert e < 1e-3


def test_fmm_converges_with_order():
    # crates/pex/tests/fmm.rs:49-60
    pts = O.lcg_points(800, 999)
    q = np.array([((i * 7) % 11) - 5.0 for i in range(800)])
    exact = O.dense_p2p(pts, q)
    e3 = rel(O.fmm_matvec(pts, q, 3, 3), exact)
    e5 = rel(O.fmm_matvec(pts, q, 5, 3), exact)
    assert e5 < e3 and e5 < 1e-4


def test_farfield_plus_near_is_full():
    pts = O.lcg_points(500, 5)
    q = np.cos(np.arange(500))
    full = O.fmm_matvec(pts, q, 4, 3, with_p2p=True)
    far = O.fmm_matvec(pts, q, 4, 3, with_p2p=False)
    # the near part is the exact sum over leaf + neighbour boxes: recompute it
    exact = O.dense_p2p(pts, q)
    assert rel(full, exact) < 1e-3
    assert np.linalg.norm(full - far) > 0.1 * np.linalg.norm(full)


def test_compressed_and_dense_m2l_agree():
    # The reference truncates the M2L CPQR where the remaining squared column
    # norm falls below eps = max(1e-3 * 0.1^n, 1e-14) times the largest one
    # (bbfmm.rs:150, linalg.rs:293-307), i.e. at sqrt(eps) in norm.  The
    # uncompressed operator must be at least as accurate against the exact
    # sum, and the two must differ by less than 1.5x the interpolation error.
    pts = O.lcg_points(900, 12345)
    q = np.array([(i % 13) - 6.0 for i in range(900)])
    exact = O.dense_p2p(pts, q)
    for n in (3, 4, 5, 6):
        a = O.fmm_matvec(pts, q, n, 3, m2l_mode=0)
        b = O.fmm_matvec(pts, q, n, 3, m2l_mode=1)
        assert rel(b, exact) <= rel(a, exact)
        assert rel(a, b) < 1.5 * rel(b, exact)
