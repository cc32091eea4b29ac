"""Oracle L2 projections (nodal normals, surface gradients) pinned by closed forms.

ComputationalDomain::compute_normals (source/computational_domain.cc:1525-1620) and
BEMProblem::compute_surface_gradients (source/bem_problem.cc:1154-1290) project onto the Q1
space with the mass matrix.  A projected field that already lies in that space comes back
exactly, which fixes the mass matrix, the right-hand side and the surface-gradient formula:
  * the outward unit normal of a flat face is constant;
  * on the faces of a box, phi = x y restricts to a bilinear (or linear) function whose
    surface gradient (y, x, 0) minus its normal part is linear in the face coordinates;
On the curved sphere the projections converge to the radial normal and to the tangential
gradient of z.  The sharp Dirichlet-Dirichlet constraint rule
(source/bem_problem.cc:1053-1078) then returns dphi/dn = a.n exactly for phi = a.x.
"""
import numpy as np
import pytest

import oracle as O
from wavebem_amd import mesh as M


def _box_exact_normals(m):
    ex = np.zeros((m.n_dofs, 3))
    for ax in range(3):
        for sd in (0, 1):
            ex[m.dof_material == M.BOX + 2 * ax + sd, ax] = 1.0 if sd else -1.0
    return ex


@pytest.mark.parametrize("normalize", [True, False])
def test_box_normals_exact(normalize):
    m = M.box_mesh(n=4, lo=(-0.3, 0.1, 0.0), hi=(0.9, 0.7, 1.3), split_top=True)
    nr = O.node_normals(m.support_points, m.cell_dofs, m.cell_orient, normalize=normalize)
    assert np.abs(nr - _box_exact_normals(m)).max() < 1e-13


def test_box_bilinear_surface_gradient_exact():
    m = M.box_mesh(n=5, lo=(-0.3, 0.1, 0.0), hi=(0.9, 0.7, 1.3))
    x, y, _ = m.support_points.T
    g = O.surface_gradients(m.support_points, m.cell_dofs, m.cell_orient, x * y)
    full = np.stack([y, x, np.zeros_like(x)], 1)
    nn = _box_exact_normals(m)
    tang = full - (full * nn).sum(1)[:, None] * nn
    assert np.abs(g - tang).max() < 1e-12


def test_sphere_projections_converge():
    """Projections onto the patches of the cube sphere: the median nodal error falls at
    (at least) second order per halving of h; the RMS error, which the one-sided patch
    edges dominate, at better than first order; every normal points outward."""
    med_n, rms_n, med_g, rms_g = [], [], [], []
    for n in (4, 8):
        m = M.cube_sphere_mesh(n=n)
        X = m.support_points
        rad = X / np.linalg.norm(X, axis=1)[:, None]
        nr = O.node_normals(X, m.cell_dofs, m.cell_orient)
        assert (nr * rad).sum(1).min() > 0.97
        g = O.surface_gradients(X, m.cell_dofs, m.cell_orient, X[:, 2])
        ez = np.array([0.0, 0.0, 1.0])
        tang = ez[None, :] - rad[:, 2:3] * rad
        en = np.linalg.norm(nr - rad, axis=1)
        eg = np.linalg.norm(g - tang, axis=1)
        med_n.append(np.median(en))
        rms_n.append(np.sqrt((en ** 2).mean()))
        med_g.append(np.median(eg))
        rms_g.append(np.sqrt((eg ** 2).mean()))
    assert med_n[0] / med_n[1] > 3.5 and med_g[0] / med_g[1] > 3.5
    assert rms_n[0] / rms_n[1] > 2.2 and rms_g[0] / rms_g[1] > 2.2
    assert med_n[1] < 0.02 and med_g[1] < 0.02


def test_sharp_dirichlet_edge_gives_exact_normal_derivative():
    m = M.box_mesh(n=4, dirichlet_faces=[(2, 1), (0, 1)], split_top=True)
    ptr, idx = M.double_nodes_csr(m.support_points)
    a = np.array([0.3, -0.7, 1.1])
    nn = _box_exact_normals(m)
    bc = np.where(m.surface_nodes == 1, m.support_points @ a, nn @ a)
    nr = O.node_normals(m.support_points, m.cell_dofs, m.cell_orient)
    gr = O.surface_gradients(m.support_points, m.cell_dofs, m.cell_orient,
                             bc * m.surface_nodes)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc, nr, gr)
    sharp = [(r, cons.inhom[l]) for l, r in enumerate(cons.rows)
             if m.surface_nodes[r] == 1 and cons.ptr[l + 1] == cons.ptr[l]]
    flat = [r for l, r in enumerate(cons.rows)
            if m.surface_nodes[r] == 1 and cons.ptr[l + 1] > cons.ptr[l]]
    assert len(sharp) == 10          # 5 edge nodes x 2 Dirichlet dofs
    assert len(flat) == 5            # the split line of the top face
    assert max(abs(v - nn[r] @ a) for r, v in sharp) < 1e-13


def test_prepare_bem_vectors_closed_forms():
    """FreeSurface::prepare_bem_vectors (source/free_surface.cc:8744-8883) restated: masks
    from the free-surface cells; on a flat wall patch the unnormalised projected normal is
    the exact unit normal, so bem_bc = -n.V exactly; the hull takes -n.V with n within the
    projection error of the analytic Wigley normal; transom dofs average the previous
    dphi_dn of their coincident dofs."""
    from tests import helpers as T

    m = T.wigley_small()
    ptr, idx = M.double_nodes_csr(m.support_points)
    V = np.array([0.9, -0.2, 0.35])
    wind = np.tile(V, (m.n_dofs, 1))
    rng = np.random.default_rng(3)
    bc0 = rng.standard_normal(m.n_dofs)
    prev = rng.standard_normal(m.n_dofs)
    inflow = rng.standard_normal(m.n_dofs)
    # transom flags on three free-surface dofs that have coincident dofs
    fs = m.dof_material == M.FREE_SURFACE
    cand = np.flatnonzero(fs & (np.diff(ptr) > 1))
    tr = np.zeros(m.n_dofs, dtype=np.uint8)
    tr[cand[:3]] = 1
    surf, oth, bc, dn = O.prepare_bem_vectors(
        m.support_points, m.cell_dofs, m.cell_orient, m.cell_material, ptr, idx,
        [M.FREE_SURFACE], [M.HULL, M.BOTTOM], [M.INFLOW], wind, bc0, bc0.copy(),
        inflow_dphi_dn=inflow, transom=tr, prev_dphi_dn=prev)
    expect_surf = m.surface_nodes.copy()
    expect_surf[tr == 1] = 0.0
    assert np.array_equal(surf, expect_surf) and np.array_equal(oth, 1.0 - expect_surf)
    bot = m.dof_material == M.BOTTOM
    assert np.abs(bc[bot] - V[2]).max() < 1e-13          # n = (0, 0, -1)
    inl = m.dof_material == M.INFLOW
    assert np.array_equal(bc[inl], inflow[inl])
    rest = np.isin(m.dof_material, [M.CENTREPLANE, M.FAR_SIDE, M.OUTFLOW])
    assert np.all(bc[rest] == 0.0) and np.all(dn[rest] == 0.0)
    hull = m.dof_material == M.HULL
    nh = M.wigley_hull_normal_out_of_fluid(m.support_points[hull])
    assert np.median(np.abs(bc[hull] + nh @ V)) < 5e-3
    free = fs & (tr == 0)
    assert np.array_equal(bc[free], bc0[free])
    for i in cand[:3]:
        dup = [d for d in idx[ptr[i]:ptr[i + 1]] if d != i]
        assert abs(bc[i] - np.mean(prev[dup])) < 1e-15 and dn[i] == bc[i]


def test_potential_gradient_closed_form_on_box():
    """phi = a.x, dphi/dn = a.n on every face: the complete gradient is a on every cell, so
    the projection returns a at every dof, before and after the double-node mean."""
    m = M.box_mesh(4, lo=(-0.2, 0.1, 0.0), hi=(0.7, 0.9, 1.3))
    ptr, idx = M.double_nodes_csr(m.support_points)
    a = np.array([0.3, -0.7, 1.1])
    phi = m.support_points @ a
    psi = O.node_normals(m.support_points, m.cell_dofs, m.cell_orient) @ a
    for avg in (False, True):
        g = O.potential_gradients(m.support_points, m.cell_dofs, m.cell_orient, phi, psi, ptr,
                                  idx, average=avg)
        assert np.abs(g - a).max() < 1e-13


def test_potential_gradient_without_psi_is_the_surface_gradient():
    """With dphi/dn = 0 the projection is the surface-gradient projection of the C oracle."""
    m = M.cube_sphere_mesh(n=4)
    ptr, idx = M.double_nodes_csr(m.support_points)
    phi = np.random.default_rng(3).standard_normal(m.n_dofs)
    g = O.potential_gradients(m.support_points, m.cell_dofs, m.cell_orient, phi,
                              np.zeros(m.n_dofs), ptr, idx, average=False)
    ref = O.surface_gradients(m.support_points, m.cell_dofs, m.cell_orient, phi)
    assert np.abs(g - ref).max() < 1e-12 * np.abs(ref).max()


def test_potential_gradient_on_sphere_converges_to_harmonic_gradient():
    """phi = z (harmonic), dphi/dn = n_z on the unit sphere: the projected complete gradient
    tends to e_z as the mesh is refined (first order at the nodes in the max norm: the flat
    cell normals carry the n psi term), and coincident dofs agree after the mean."""
    errs = []
    for k in (4, 8):
        m = M.cube_sphere_mesh(n=k)
        ptr, idx = M.double_nodes_csr(m.support_points)
        X = m.support_points
        nrm = X / np.linalg.norm(X, axis=1)[:, None]
        g = O.potential_gradients(X, m.cell_dofs, m.cell_orient, X[:, 2], nrm[:, 2], ptr, idx)
        errs.append(np.abs(g - np.array([0.0, 0.0, 1.0])).max())
        for i in range(m.n_dofs):
            assert np.array_equal(g[i], g[idx[ptr[i]]])
    assert errs[0] < 0.05 and errs[1] < errs[0] / 1.8, errs
