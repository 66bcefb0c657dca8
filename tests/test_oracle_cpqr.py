"""Pins the oracle's LowRankApprox::compress (linalg.rs:277-381) and its
M2L truncation rule (bbfmm.rs:132-174), and the 7-point triangle rule of
field_integral (integrals/panel.rs:213-253), by what the mathematics fixes:
exact-rank matrices, diagonal matrices whose cut-off only the squared-norm
rule gives, an independent pivoted QR (LAPACK dgeqp3 through scipy), the
closed-form Dunavant/Strang-Fix points and weights, and exact monomial
integrals on the reference triangle."""
import math

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as O


def test_exact_rank_matrix_returns_its_rank():
    rng = np.random.default_rng(1)
    for m, n, r in ((40, 30, 7), (25, 60, 12), (64, 64, 1)):
        a = rng.normal(size=(m, r)) @ rng.normal(size=(r, n))
        # the downdated squared norms of the dependent columns end at rounding
        # level (~1e-16 of the largest), below eps = 1e-12
        u, vt, k = O.lowrank_compress(a, 1e-12)
        assert k == r
        assert np.abs(u @ vt - a).max() < 1e-13 * np.abs(a).max()
        assert np.abs(u.T @ u - np.eye(k)).max() < 1e-13


def test_cutoff_is_on_squared_column_norms():
    # squared norms 1, .25, 2.5e-3, 2.5e-5, 2.5e-7 against eps = 1e-4: three
    # columns stay.  A rule on the norms themselves would keep five, a rule on
    # sqrt(eps) two.
    d = np.array([1.0, 0.5, 0.05, 0.005, 0.0005])
    u, vt, k = O.lowrank_compress(np.diag(d), 1e-4)
    assert k == 3
    assert np.array_equal(u @ vt, np.diag(np.r_[d[:3], 0.0, 0.0]))


def test_threshold_tie_is_kept():
    # squared norm 2^-14 exactly equal to eps * max: kept (the loop stops
    # only below the threshold, linalg.rs:303)
    d = np.array([1.0, 2.0 ** -7])
    _, _, k = O.lowrank_compress(np.diag(d), 2.0 ** -14)
    assert k == 2
    _, _, k = O.lowrank_compress(np.diag(d), 2.0 ** -14 * 1.0000001)
    assert k == 1


def test_pivots_take_the_largest_remaining_column_first():
    d = np.array([0.05, 1.0, 0.0005, 0.5, 0.005])
    u, vt, k = O.lowrank_compress(np.diag(d), 1e-4)
    assert k == 3
    # U's columns are the unit vectors of the pivots in order 1, 3, 0
    assert np.array_equal(u, np.eye(5)[:, [1, 3, 0]])
    assert np.array_equal(vt, np.diag(d)[[1, 3, 0]])


def _lapack_truncation(a, eps):
    """Pivoted Householder QR (dgeqp3): rank = leading |R_ss|^2 at or above
    eps * max squared column norm; the truncated product Q_k R_k P^T."""
    q, r, piv = sl.qr(a, pivoting=True, mode="economic")
    thr = eps * (a * a).sum(axis=0).max()
    d = np.abs(np.diag(r)) ** 2
    k = int(np.argmax(d < thr)) if np.any(d < thr) else len(d)
    out = np.zeros_like(a)
    out[:, piv] = q[:, :k] @ r[:k, :]
    return k, out


def test_matches_lapack_pivoted_qr_on_decaying_spectra():
    rng = np.random.default_rng(7)
    for m, n in ((50, 40), (64, 64), (30, 45)):
        qa, _ = np.linalg.qr(rng.normal(size=(m, m)))
        qb, _ = np.linalg.qr(rng.normal(size=(n, n)))
        s = 10.0 ** (-np.arange(min(m, n)) / 4.0)
        a = qa[:, :len(s)] @ np.diag(s) @ qb[:, :len(s)].T
        for eps in (1e-6, 1e-9):
            u, vt, k = O.lowrank_compress(a, eps)
            kl, al = _lapack_truncation(a, eps)
            assert k == kl
            assert np.abs(u @ vt - al).max() < 1e-12 * np.abs(a).max()


@pytest.mark.parametrize("order", [3, 4, 5])
@pytest.mark.parametrize("off", [(2, 1, 3), (3, -2, 1), (-3, 2, -1)])
def test_m2l_truncation_matches_lapack(order, off):
    # offsets with no reflection or axis-swap symmetry: no exactly tied
    # column norms, so both QRs pick the same pivots
    g = O.m2l_matrix(order, off)
    eps = O.m2l_eps(order)
    assert eps == max(1e-3 * 0.1 ** order, 1e-14)
    u, vt, k = O.lowrank_compress(g, eps)
    kl, gl = _lapack_truncation(g, eps)
    assert k == kl and 0 < k < order ** 3
    assert np.abs(u @ vt - gl).max() < 1e-12 * np.abs(g).max()


def test_m2l_matrix_is_the_kernel_on_chebyshev_nodes():
    n = 3
    nd = np.cos((2 * np.arange(n) + 1) * np.pi / (2 * n))
    X = np.array([(a, b, c) for a in nd for b in nd for c in nd])
    off = np.array([2, -3, 0])
    d = X[:, None, :] - X[None, :, :] - 2 * off
    assert np.allclose(O.m2l_matrix(n, off), 1 / np.linalg.norm(d, axis=2), rtol=1e-15, atol=0)


# ---------------------------------------------------------------- TRI7

S15 = math.sqrt(15.0)
# points (1 - 2a, a, a) and permutations with a = (6 +- sqrt 15) / 21
A1, S1 = (6 + S15) / 21, (9 - 2 * S15) / 21   # 0.4701..., 0.0597...
A2, S2 = (6 - S15) / 21, (9 + 2 * S15) / 21   # 0.1012..., 0.7974...
W0, W1, W2 = 9 / 40, (155 + S15) / 1200, (155 - S15) / 1200
EXACT = np.array([[1 / 3, 1 / 3, 1 / 3, W0],
                  [S1, A1, A1, W1], [A1, S1, A1, W1], [A1, A1, S1, W1],
                  [S2, A2, A2, W2], [A2, S2, A2, W2], [A2, A2, S2, W2]])


def test_tri7_table_is_the_degree5_rule():
    # the seven points and weights of the degree-5 rule in closed form; the
    # reference's table carries 15 digits
    t = O.tri7()
    assert np.abs(t - EXACT).max() < 1e-14
    assert np.all(np.abs(t[:, :3].sum(axis=1) - 1) < 1e-14)


def test_tri7_integrates_monomials_to_degree5_exactly():
    t = O.tri7()
    x, y, w = t[:, 1], t[:, 2], t[:, 3]
    for a in range(6):
        for b in range(6 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs(0.5 * (w * x ** a * y ** b).sum() - exact) < 1e-15
    # and not degree 6: the rule is no better than it claims
    assert abs(0.5 * (w * x ** 6).sum() - math.factorial(6) / math.factorial(8)) > 1e-6


def _field_rule(verts, p, n_obs):
    """The 7-point normal-field sum over the fan triangles, closed-form points."""
    v = np.asarray(verts, float).reshape(-1, 3)
    acc = 0.0
    for i in range(1, len(v) - 1):
        v0, v1, v2 = v[0], v[i], v[i + 1]
        ar = 0.5 * np.linalg.norm(np.cross(v1 - v0, v2 - v0))
        if ar == 0:
            continue
        for l0, l1, l2, w in EXACT:
            d = l0 * v0 + l1 * v1 + l2 * v2 - p
            r2 = d @ d
            if r2 > 0:
                acc += w * ar * (d @ n_obs) / (r2 * math.sqrt(r2))
    return acc


def test_field_integral_is_the_seven_point_rule():
    rng = np.random.default_rng(3)
    quad = np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0]], float)
    tri = np.array([[0, 0, 0], [1.5, 0.2, 0.1], [0.3, 1.1, -0.2]], float)
    for verts in (quad, tri):
        for _ in range(20):
            p = rng.normal(size=3) * 3
            n_obs = rng.normal(size=3)
            n_obs /= np.linalg.norm(n_obs)
            got = O.field_integral(verts, p, n_obs)
            want = _field_rule(verts, p, n_obs)
            assert abs(got - want) <= 1e-13 * max(1e-30, abs(want)) + 1e-16
