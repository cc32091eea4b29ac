"""Adjoint of the BEM solve and the hull drag functional (verification harness).

The BEM solve u(x) of ``BEMProblem.solve_system`` becomes a ``torch.autograd.Function`` of
the support points x: the forward moves the mesh (``wb_set_support_points``), assembles and
solves on the GPU; the backward solves the transposed constrained system
``A_c^T lambda = dJ/du`` (``wb_solve_adjoint``) and applies the backward of the assembly
(``wb_residual_vjp``): dJ/dx = dJ/dx|_u - lambda^T dR/dx.  A C++ driver calls the same two
entry points.

The drag functional is the linearised Bernoulli pressure integrated over the hull,
``D = sum_cells sum_q p_q (n JxW)_x`` with ``p = -rho (phi_t + U dphi/dx)`` and
``grad phi = surface gradient of the Q1 interpolant of phi + (dphi/dn) n``, on the 4x4
Gauss points of each hull cell.  It is written in torch (float64), so its own partial
derivatives come from autograd.  The reference evaluates the hull pressure in
FreeSurface::compute_pressure (source/free_surface.cc:9407-9725) from the DAE state; this
functional keeps its linearised Bernoulli form for a single BEM state.
"""
from __future__ import annotations

import numpy as np
import torch

RHO = 1025.1  # source/free_surface.cc:9417


def _gauss01(q):
    t, w = np.polynomial.legendre.leggauss(q)
    x = 0.5 * (t + 1.0)
    w = 0.5 * w
    xi = np.tile(x, q)            # point i + q*j: xi fastest
    eta = np.repeat(x, q)
    ww = np.repeat(w, q) * np.tile(w, q)
    return xi, eta, ww


def hull_drag(support, cells, orient, phi, dphi_dn, phi_t, U, rho=RHO, q_order=4):
    """Linearised Bernoulli drag over ``cells`` (k x 4 dof indices, deal.II vertex order,
    ``orient`` = +-1 so that the normal points out of the fluid).  All tensor arguments may
    require grad; returns a 0-d float64 tensor."""
    xi, eta, w = (torch.as_tensor(a, dtype=torch.float64) for a in _gauss01(q_order))
    sh = torch.stack([(1 - xi) * (1 - eta), xi * (1 - eta), (1 - xi) * eta, xi * eta], 1)
    dx = torch.stack([-(1 - eta), (1 - eta), -eta, eta], 1)
    de = torch.stack([-(1 - xi), -xi, (1 - xi), xi], 1)
    cells = torch.as_tensor(np.asarray(cells), dtype=torch.long)
    o = torch.as_tensor(np.asarray(orient), dtype=torch.float64)
    V = support[cells]                                    # (k, 4, 3)
    a = torch.einsum("qv,cvd->cqd", dx, V)
    b = torch.einsum("qv,cvd->cqd", de, V)
    cr = torch.cross(a, b, dim=2)
    cl = torch.linalg.norm(cr, dim=2)
    n = o[:, None, None] * cr / cl[:, :, None]
    N = o[:, None, None] * w[None, :, None] * cr          # n JxW
    pc = phi[cells]                                       # (k, 4)
    dphi_xi = pc @ dx.T                                   # (k, q)
    dphi_eta = pc @ de.T
    g00 = (a * a).sum(2)
    g01 = (a * b).sum(2)
    g11 = (b * b).sum(2)
    det = g00 * g11 - g01 * g01
    c_a = (g11 * dphi_xi - g01 * dphi_eta) / det
    c_b = (-g01 * dphi_xi + g00 * dphi_eta) / det
    psi_q = dphi_dn[cells] @ sh.T
    grad_x = c_a * a[:, :, 0] + c_b * b[:, :, 0] + psi_q * n[:, :, 0]
    pt_q = phi_t[cells] @ sh.T
    p = -rho * (pt_q + U * grad_x)
    return (p * N[:, :, 0]).sum()


class _BEMSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, support, problem, tmp_rhs, release):
        x = np.ascontiguousarray(support.detach().cpu().numpy(), dtype=np.float64)
        problem.set_support_points(x)
        problem.assemble_system()
        phi = tmp_rhs * problem.surface_nodes
        dphi = tmp_rhs * problem.other_nodes
        problem.solve_system(phi, dphi, tmp_rhs)
        if release:
            # the backward assembles G and H again at this geometry
            problem.release_matrices()
        ctx.problem = problem
        ctx.x = x
        ctx.phi, ctx.dphi = phi, dphi
        return torch.from_numpy(phi.copy()), torch.from_numpy(dphi.copy())

    @staticmethod
    def backward(ctx, g_phi, g_dphi):
        p = ctx.problem
        # the backward needs the operator assembled at the forward's geometry
        p.set_support_points(ctx.x)
        p.assemble_system()
        g_u = (g_phi.detach().numpy() * p.other_nodes
               + g_dphi.detach().numpy() * p.surface_nodes)
        lam = p.solve_adjoint(g_u)
        grad = -p.residual_vjp(lam, ctx.phi, ctx.dphi, free_rows_only=True)
        return torch.from_numpy(grad), None, None, None


def bem_solve(support, problem, tmp_rhs, release=False):
    """(phi, dphi_dn) of solve_system at support points ``support`` ((n, 3) float64 tensor),
    differentiable with respect to ``support`` (the boundary data ``tmp_rhs`` and the
    installed masks and constraints are held fixed).  The backward always assembles G and H
    again at the forward's geometry; with ``release`` the forward frees them once it has
    solved, so the device memory is free between the forward and the backward."""
    return _BEMSolve.apply(support, problem, np.ascontiguousarray(tmp_rhs, dtype=np.float64),
                           bool(release))


def bem_solve_jvp(problem, support, tmp_rhs, dx):
    """Tangent of the BEM solve: (phi, dphi_dn) at ``support`` and their directional
    derivatives along a support-point direction ``dx`` (n, 3), boundary data fixed:
    du = -A_c^-1 (dR/dx dx), scattered like solve_system (phi on Neumann dofs, dphi_dn on
    Dirichlet dofs; the known entries do not move)."""
    x = np.ascontiguousarray(support, dtype=np.float64)
    tmp = np.ascontiguousarray(tmp_rhs, dtype=np.float64)
    problem.set_support_points(x)
    problem.assemble_system()
    phi = tmp * problem.surface_nodes
    dphi = tmp * problem.other_nodes
    problem.solve_system(phi, dphi, tmp)
    dr = problem.residual_jvp(dx, phi, dphi, free_rows_only=True)
    du = problem.solve_rhs(-dr)
    return phi, dphi, du * problem.other_nodes, du * problem.surface_nodes


def class_gradient(grad, dn_ptr, dn_idx):
    """Per double-node class: moving a geometric node moves all its coincident dofs, so
    the gradient of a node is the sum over its dofs.  Returns (n, 3) with every dof holding
    its class total."""
    g = np.asarray(grad)
    out = np.empty_like(g)
    for i in range(g.shape[0]):
        out[i] = g[dn_idx[dn_ptr[i]:dn_ptr[i + 1]]].sum(axis=0)
    return out
