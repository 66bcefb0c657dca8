"""The library's truncated M2L operators (gp_m2l_operator, host code of
gp_m2l_cpqr.cpp) against the oracle's LowRankApprox::compress of the same
matrices (bbfmm.rs:132-174, linalg.rs:277-381): the same rank for every
far-field offset and the same product U V^T.  No device needed."""
import itertools

import numpy as np
import pytest

import gpurify_amd as G
from oracle import oracle as O

FAR = [o for o in itertools.product(range(-3, 4), repeat=3) if max(map(abs, o)) > 1]


def test_far_offsets_are_316():
    assert len(FAR) == 316


@pytest.mark.parametrize("order", [3, 4])
def test_every_offset_matches_the_oracle(order):
    eps = O.m2l_eps(order)
    for off in FAR:
        g = O.m2l_matrix(order, off)
        u, vt, k = O.lowrank_compress(g, eps)
        gc, kc = G.m2l_operator(order, off, compressed=True)
        assert kc == k, off
        assert np.abs(gc - u @ vt).max() <= 1e-14 * np.abs(g).max(), off
        gu, ku = G.m2l_operator(order, off, compressed=False)
        assert ku == order ** 3 and np.array_equal(gu, g)


@pytest.mark.parametrize("order", [5, 6])
def test_sampled_offsets_match_the_oracle(order):
    eps = O.m2l_eps(order)
    for off in FAR[::23]:
        g = O.m2l_matrix(order, off)
        u, vt, k = O.lowrank_compress(g, eps)
        gc, kc = G.m2l_operator(order, off, compressed=True)
        assert kc == k, off
        assert np.abs(gc - u @ vt).max() <= 1e-14 * np.abs(g).max(), off


def test_truncation_is_below_the_interpolation_scale():
    # the truncated operator stays within sqrt(eps) (in norm) of the full one
    for order in (3, 4, 5, 6):
        g, _ = G.m2l_operator(order, (2, 0, 0), compressed=False)
        gc, k = G.m2l_operator(order, (2, 0, 0), compressed=True)
        assert 0 < k < order ** 3
        assert np.linalg.norm(gc - g) <= 2 * np.sqrt(O.m2l_eps(order)) * np.linalg.norm(g) * order ** 1.5


def test_rejects_near_and_out_of_range_offsets():
    for off in ((1, 0, 0), (0, 0, 0), (4, 0, 0)):
        with pytest.raises(G.GpError):
            G.m2l_operator(4, off)
    with pytest.raises(G.GpError):
        G.m2l_operator(7, (2, 0, 0))


@pytest.mark.parametrize("order", [3, 4])
def test_factors_are_the_references(order):
    # gp_m2l_factors: the u and vt of LowRankApprox::compress themselves (the
    # FP64 operator's factored M2L kernel applies u (vt x), as bbfmm.rs does)
    eps = O.m2l_eps(order)
    for off in FAR:
        u, vt, k = O.lowrank_compress(O.m2l_matrix(order, off), eps)
        fu, fvt, fk = G.m2l_factors(order, off)
        assert fk == k, off
        assert np.abs(fu - u).max() <= 1e-15 and np.abs(fvt - vt).max() <= 1e-14 * np.abs(vt).max()
        gc, _ = G.m2l_operator(order, off, compressed=True)
        assert np.abs(fu @ fvt - gc).max() <= 1e-14 * np.abs(gc).max(), off
    with pytest.raises(G.GpError):
        G.m2l_factors(order, (1, 1, 0))
