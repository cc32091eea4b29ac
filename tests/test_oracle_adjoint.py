"""Oracle backward of the assembly (orc_residual_vjp) pinned by central differences of the
oracle's own assembly, and the drag functional's autograd against finite differences."""
import numpy as np
import pytest
import torch

import oracle as O
from wavebem_amd import mesh as M
from tests import helpers as T
from wavebem_amd.adjoint import hull_drag


def _fd_check(m, seed, n_dofs=5, h=1e-6):
    ptr, idx = M.double_nodes_csr(m.support_points)
    n = m.n_dofs
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(n)
    phi = rng.standard_normal(n)
    psi = rng.standard_normal(n)
    X = m.support_points.copy()

    def F(Xs):
        H, G = O.assemble(Xs, m.cell_dofs, m.cell_orient, ptr, idx)
        return w @ (H @ phi - H.sum(1) * phi - G @ psi)

    g = O.residual_vjp(X, m.cell_dofs, m.cell_orient, ptr, idx, w, phi, psi)
    fd, an = [], []
    for d in rng.choice(n, n_dofs, replace=False):
        for k in range(3):
            Xp, Xm = X.copy(), X.copy()
            Xp[d, k] += h
            Xm[d, k] -= h
            fd.append((F(Xp) - F(Xm)) / (2 * h))
            an.append(g[d, k])
    fd, an = np.array(fd), np.array(an)
    return np.abs(fd - an).max() / np.abs(fd).max()


@pytest.mark.parametrize("kind", ["box", "sphere"])
def test_residual_vjp_matches_central_differences(kind):
    """Flat cells (box, singular cells on both sides of each edge) and curved cells
    (sphere, where N and J change with every vertex)."""
    m = M.box_mesh(n=3, dirichlet_top=True) if kind == "box" else M.cube_sphere_mesh(n=2)
    err = _fd_check(m, seed=3)
    print(f"  {kind}: max |vjp - fd| / max |fd| = {err:.2e}")
    assert err < 1e-7


def test_residual_vjp_is_linear_in_weights():
    m = M.cube_sphere_mesh(n=2)
    ptr, idx = M.double_nodes_csr(m.support_points)
    rng = np.random.default_rng(4)
    n = m.n_dofs
    w1, w2, phi, psi = (rng.standard_normal(n) for _ in range(4))
    args = (m.support_points, m.cell_dofs, m.cell_orient, ptr, idx)
    g1 = O.residual_vjp(*args, w1, phi, psi)
    g2 = O.residual_vjp(*args, w2, phi, psi)
    g12 = O.residual_vjp(*args, 2.0 * w1 - 3.0 * w2, phi, psi)
    assert np.abs(g12 - (2 * g1 - 3 * g2)).max() <= 1e-12 * np.abs(g12).max()
    # constant phi and zero psi: R = H (phi_j - phi_i) vanishes for every geometry
    g0 = O.residual_vjp(*args, w1, np.full(n, 2.5), np.zeros(n))
    assert np.abs(g0).max() < 1e-12


def test_hull_drag_autograd_matches_central_differences():
    m = T.wigley_small()
    hull = np.nonzero(m.cell_material == M.HULL)[0]
    cells, orient = m.cell_dofs[hull], m.cell_orient[hull]
    rng = np.random.default_rng(5)
    n = m.n_dofs
    phi = torch.tensor(rng.standard_normal(n) * 1e-2, dtype=torch.float64, requires_grad=True)
    dphi = torch.tensor(rng.standard_normal(n) * 1e-2, dtype=torch.float64, requires_grad=True)
    pt = torch.tensor(rng.standard_normal(n) * 1e-3, dtype=torch.float64)
    X = torch.tensor(m.support_points, dtype=torch.float64, requires_grad=True)
    J = hull_drag(X, cells, orient, phi, dphi, pt, U=1.0)
    J.backward()
    hull_dofs = np.unique(cells)
    h = 1e-6
    errs = []
    for d in rng.choice(hull_dofs, 4, replace=False):
        for k in range(3):
            Xp = X.detach().clone()
            Xm = X.detach().clone()
            Xp[d, k] += h
            Xm[d, k] -= h
            fd = (hull_drag(Xp, cells, orient, phi.detach(), dphi.detach(), pt, U=1.0)
                  - hull_drag(Xm, cells, orient, phi.detach(), dphi.detach(), pt, U=1.0)) / (2 * h)
            errs.append((float(fd), float(X.grad[d, k])))
    e = np.array(errs)
    assert np.abs(e[:, 0] - e[:, 1]).max() <= 1e-6 * np.abs(e[:, 0]).max()
    # linear in phi: the phi-gradient is the functional's coefficient vector
    phi2 = phi.detach() * 2.0
    J2 = hull_drag(X.detach(), cells, orient, phi2, dphi.detach(), pt, U=1.0)
    J0 = hull_drag(X.detach(), cells, orient, torch.zeros_like(phi2), dphi.detach(), pt, U=1.0)
    assert abs(float(J2 - J0) - 2.0 * float(phi.grad @ phi.detach())) < 1e-12 * max(1.0, abs(float(J2)))
