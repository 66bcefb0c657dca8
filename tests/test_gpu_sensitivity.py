"""Adjoint sensitivities dC/dtheta (gp_cap_jacobian) on the device, pinned
by central finite differences of the device solves and of the oracle, and by
what the mathematics fixes: C scales with the geometry (C(lambda x) =
lambda C, so dC/dlambda = C) and does not change when everything translates."""
import numpy as np
import pytest

import gpurify_amd as G
from gpurify_amd import sensitivity as S
from gpurify_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def wires():
    v, c, _ = G.mesh_groups(W.two_wires(), refine=1)  # 173 panels: dense
    params, names = S.wire_params(v, c, [0, 1])
    return v, c, params, names


def test_dense_jacobian_matches_device_fd(ctx, wires):
    v, c, params, names = wires
    o = G.default_opts(method=G.GP_METHOD_DENSE)
    C0, J, rep = ctx.cap_jacobian(v, c, 3, params, o)
    assert rep["method_used"] == G.GP_METHOD_DENSE and J.shape == (6, 3, 3)
    fd = S.central_fd(lambda vv: ctx.cap_solve(vv, c, 3, o)[0], v, params, 1e-10)
    for p, name in enumerate(names):
        assert rel(J[p], fd[p]) < 1e-6, name
    # width grows every |C_ij| of its wire; spacing moves charge between neighbours
    assert J[0][0, 0] > 0 and J[3][1, 1] > 0


def test_dense_jacobian_matches_oracle_fd(ctx, wires):
    v, c, params, names = wires
    _, J, _ = ctx.cap_jacobian(v, c, 3, params, G.default_opts(method=G.GP_METHOD_DENSE))
    fd = S.central_fd(lambda vv: O.cap_dense(vv, c, 3), v, params, 1e-10)
    for p, name in enumerate(names):
        assert rel(J[p], fd[p]) < 1e-6, name


def test_scale_covariance_and_translation_invariance(ctx, wires):
    v, c, _, _ = wires
    n = len(v)
    allp = np.arange(n, dtype=np.uint32)
    centre = v.reshape(-1, 3).mean(0)
    dil = (v - centre)                       # d v / d lambda at lambda = 1
    shift = np.zeros_like(v)
    shift[:, :, 0] = 1.0                     # rigid translation in x
    lift = np.zeros_like(v)
    lift[:, :, 2] = 1.0                      # rigid translation in z (plate too)
    C0, J, _ = ctx.cap_jacobian(v, c, 3, [(allp, dil), (allp, shift), (allp, lift)],
                                G.default_opts(method=G.GP_METHOD_DENSE))
    assert rel(J[0], C0) < 1e-9              # dC/dlambda = C
    L = np.ptp(v.reshape(-1, 3), axis=0).max()
    assert np.abs(J[1]).max() * L < 1e-9 * np.abs(C0).max()
    assert np.abs(J[2]).max() * L < 1e-9 * np.abs(C0).max()


def test_fmm_jacobian_matches_dense_and_fd(ctx):
    v, c, _ = G.mesh_groups(W.two_wires(), refine=3)  # 1557 panels
    params, names = S.wire_params(v, c, [0, 1])
    _, Jd, _ = ctx.cap_jacobian(v, c, 3, params, G.default_opts(method=G.GP_METHOD_DENSE,
                                                                dense_limit=5000))
    of = G.default_opts(method=G.GP_METHOD_FMM, tol=1e-12)
    _, Jf, rep = ctx.cap_jacobian(v, c, 3, params, of)
    assert rep["method_used"] == G.GP_METHOD_FMM and rep["converged"] == 1
    fd = S.central_fd(lambda vv: ctx.cap_solve(vv, c, 3, of)[0], v, params, 1e-10)
    for p, name in enumerate(names):
        # The adjoint differentiates the exact integrals; its FMM forward and
        # adjoint solutions carry the interpolation error (order 4: ~1e-4 in
        # C).  Differences of FMM solves also differentiate that error, which
        # moves with the panels: a few 1e-3.
        assert rel(Jf[p], Jd[p]) < 2e-3, name
        assert rel(Jf[p], fd[p]) < 1e-2, name
    om = G.default_opts(method=G.GP_METHOD_FMM, tol=1e-12, precision=G.GP_PREC_MIXED)
    _, Jm, rm = ctx.cap_jacobian(v, c, 3, params, om)
    assert rm["converged"] == 1 and rel(Jm, Jf) < 1e-6


def test_autograd_function_backward(ctx, wires):
    torch = pytest.importorskip("torch")
    v, c, params, _ = wires
    o = G.default_opts(method=G.GP_METHOD_DENSE)
    theta = torch.zeros(len(params), dtype=torch.float64, requires_grad=True)
    k = torch.tensor(3.9, dtype=torch.float64, requires_grad=True)
    W_ = torch.randn(3, 3, dtype=torch.float64, generator=torch.Generator().manual_seed(0))
    Cv = S.CapacitanceFunction.apply(theta, ctx, v, c, 3, params, o)
    loss = (W_ * (k * Cv)).sum()
    loss.backward()
    C0, J, _ = ctx.cap_jacobian(v, c, 3, params, o)
    want = 3.9 * np.einsum("pij,ij->p", J, W_.numpy())
    np.testing.assert_allclose(theta.grad.numpy(), want, rtol=1e-12)
    assert float(k.grad) == pytest.approx(float((W_.numpy() * C0).sum()), rel=1e-12)  # dC/dk = C/k


def test_many_conductor_jacobian(ctx):
    # 21 conductors: the mixed adjoint's transposed near field takes the
    # matrix-core kernel (blocks of more than 16 RHS) and the sensitivity
    # tiles contract 21 sigma columns (12 thread groups); mixed against FP64
    # on the same FMM operator, and the FMM sensitivities (far rules beyond
    # 6 and 30 diagonals, closed forms inside) against the dense path's
    # closed forms everywhere
    v, c, _ = G.mesh_groups(W.crossing_bundle(n=10, length=5000), refine=2)
    nc = int(c.max()) + 1
    params, names = S.wire_params(v, c, [0, 5, 12])
    of = G.default_opts(method=G.GP_METHOD_FMM, tol=1e-11)
    C64, J64, r64 = ctx.cap_jacobian(v, c, nc, params, of)
    om = G.default_opts(method=G.GP_METHOD_FMM, tol=1e-11, precision=G.GP_PREC_MIXED)
    Cm, Jm, rm = ctx.cap_jacobian(v, c, nc, params, om)
    assert r64["method_used"] == G.GP_METHOD_FMM and rm["converged"] == 1
    assert rel(Cm, C64) < 1e-8 and rel(Jm, J64) < 1e-6
    od = G.default_opts(method=G.GP_METHOD_DENSE, dense_limit=20000)
    Cd, Jd, _ = ctx.cap_jacobian(v, c, nc, params, od)
    # FMM order 4 (~1e-4 in C) and the far rules put the derivatives of the
    # weak couplings (down to 1e-4 of the diagonal here) a few percent off
    # and the whole tensor ~5e-3 (measured 5.5e-3, worst parameter 1.5e-2);
    # the mistakes this guards against — a near entry lost or counted twice,
    # a sign, a row tile dropped — move J by O(1)
    assert rel(J64, Jd) < 1e-2
    for p, name in enumerate(names):
        assert rel(J64[p], Jd[p]) < 3e-2, name
