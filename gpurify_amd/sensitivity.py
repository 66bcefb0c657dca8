"""Geometric parameters of straight wires and the forward/backward pair for
capacitance sensitivities.

A parameter is a list of (panel index, vertex velocity d vertex / d theta)
over the panels it moves; every other panel stays, so the mesh topology is
frozen.  For a straight wire (one rectangle extruded along its axis):

* width   -- the two side faces move apart, +-1/2 per unit theta, and the
             top, bottom and end faces stretch with them (cross-axis
             velocity (c - centre) / width);
* spacing -- the wire translates across its axis (velocity 1), i.e. the
             gap to its neighbours changes;
* height  -- the wire translates up (velocity 1 in z), thickness fixed.

The homogeneous dielectric k multiplies C (bridge.rs:405-407), so dC/dk = C/k
needs no solve; CapacitanceFunction leaves it to ordinary autograd.

``CapacitanceFunction`` follows torch.autograd.Function: forward returns the
vacuum Maxwell matrix at theta (vertices moved by sum_p theta_p dv_p) and
keeps the Jacobian from gp_cap_jacobian; backward returns
dL/dtheta_p = sum_ij dL/dC_ij J[p][i][j].
"""
from __future__ import annotations

import numpy as np

KINDS = ("width", "spacing", "height")


def wire_params(verts4, cond, conductors, kinds=KINDS):
    """Parameters of straight-wire conductors.  Returns (params, names) with
    params[p] = (panel indices, dverts m x 4 x 3 in metres per metre)."""
    v = np.asarray(verts4, dtype=np.float64).reshape(-1, 4, 3)
    cond = np.asarray(cond)
    params, names = [], []
    for c in conductors:
        idx = np.nonzero(cond == c)[0]
        if len(idx) == 0:
            raise ValueError(f"conductor {c} has no panels")
        pts = v[idx].reshape(-1, 3)
        lo, hi = pts.min(0), pts.max(0)
        ext = hi - lo
        axis = int(np.argmax(ext[:2]))  # along x (0) or y (1)
        cross = 1 - axis
        width = ext[cross]
        centre = 0.5 * (lo[cross] + hi[cross])
        for kind in kinds:
            dv = np.zeros((len(idx), 4, 3))
            if kind == "width":
                dv[:, :, cross] = (v[idx][:, :, cross] - centre) / width
            elif kind == "spacing":
                dv[:, :, cross] = 1.0
            elif kind == "height":
                dv[:, :, 2] = 1.0
            else:
                raise ValueError(kind)
            params.append((idx.astype(np.uint32), dv))
            names.append((int(c), kind))
    return params, names


def move(verts4, params, theta):
    """Vertices at theta: verts + sum_p theta_p dv_p."""
    v = np.array(verts4, dtype=np.float64).reshape(-1, 4, 3)
    for (idx, dv), t in zip(params, theta):
        if t != 0.0:
            v[idx] += t * np.asarray(dv).reshape(-1, 4, 3)
    return v


def central_fd(solve, verts4, params, h):
    """dC/dtheta_p by central differences of solve(verts) -> C, step h."""
    out = []
    for p in range(len(params)):
        th = np.zeros(len(params))
        th[p] = h
        cp = solve(move(verts4, params, th))
        cm = solve(move(verts4, params, -th))
        out.append((cp - cm) / (2.0 * h))
    return np.array(out)


try:
    import torch

    class CapacitanceFunction(torch.autograd.Function):
        """C(theta) with the adjoint Jacobian as its backward.

        apply(theta, ctx_dev, verts4, cond, ncond, params, opts) -> C (torch,
        float64, farads, vacuum)."""

        @staticmethod
        def forward(ctx, theta, dev, verts4, cond, ncond, params, opts):
            th = theta.detach().cpu().numpy().astype(np.float64)
            C, J, _ = dev.cap_jacobian(move(verts4, params, th), cond, ncond, params, opts)
            ctx.save_for_backward(torch.from_numpy(J))
            return torch.from_numpy(C)

        @staticmethod
        def backward(ctx, grad_c):
            (J,) = ctx.saved_tensors
            g = J.reshape(J.shape[0], -1) @ grad_c.reshape(-1).to(J.dtype)
            return g, None, None, None, None, None, None
except ImportError:  # torch is optional for the C ABI users
    CapacitanceFunction = None
