"""GPU adjoint: the transposed solve, the backward of the assembly, and drag gradients.

Bars: the backward of the assembly within 1e-11 (norm-wise) of the oracle's, itself pinned
by central differences; the transposed solve (FP64 and FP32-IR) within 1e-8 of a dense
solve; adjoint drag
gradients within 1e-6 (relative to the largest checked entry) of central differences of
the GPU forward solve with h = 1e-6.
"""
import numpy as np
import pytest
import torch

import oracle as O
from tests import helpers as T
from wavebem_amd import mesh as M
from wavebem_amd.adjoint import bem_solve, hull_drag
from wavebem_amd.bem import BEMProblem, Config, WaveBemError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["wigley_small", "box6", "sphere4"])
def test_residual_vjp_matches_oracle(gpu_available, kind):
    m, ptr, idx, H, G = T.oracle_assembly(kind)
    n = m.n_dofs
    rng = np.random.default_rng(11)
    w, phi, psi = (rng.standard_normal(n) for _ in range(3))
    o = O.residual_vjp(m.support_points, m.cell_dofs, m.cell_orient, ptr, idx, w, phi, psi)
    with BEMProblem(m, Config(), double_nodes=(ptr, idx)) as p:
        p.assemble_system()
        g = p.residual_vjp(w, phi, psi, free_rows_only=False)
    err = np.abs(g - o).max() / np.abs(o).max()
    print(f"  [{kind}] n={n} vjp rel err {err:.2e}")
    assert err < 1e-11


def test_adjoint_solve_matches_dense(gpu_available):
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    bc, _ = M.wigley_bc(m, seed=3)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    A, _ = T.explicit_system(H, G, m.surface_nodes, m.other_nodes, cons, bc)
    rng = np.random.default_rng(12)
    g = rng.standard_normal(m.n_dofs)
    lam_ref = np.linalg.solve(A.T, g)
    with BEMProblem(m, Config(solver_tolerance=1e-13), double_nodes=(ptr, idx)) as p:
        p.compute_constraints(bc)
        p.assemble_system()
        lam = p.solve_adjoint(g)
        st = p.last_stats
    err = np.abs(lam - lam_ref).max() / np.abs(lam_ref).max()
    print(f"  adjoint: {st['iterations']} it, rel err {err:.2e}")
    assert err < 1e-8


def test_adjoint_fp32_ir_matches_dense(gpu_available):
    """The transposed solve of an FP32-IR context (refinement residuals with the transposed
    hi + lo operator, inner GMRES on the transposed FP32 hi part) against a dense solve of
    the oracle's FP64 system: the hi + lo storage is within 2^-33 of the FP64 entries, so
    the bar is the FP64 one."""
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    bc, _ = M.wigley_bc(m, seed=3)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    A, _ = T.explicit_system(H, G, m.surface_nodes, m.other_nodes, cons, bc)
    rng = np.random.default_rng(12)
    g = rng.standard_normal(m.n_dofs)
    lam_ref = np.linalg.solve(A.T, g)
    # 1e-12: the refinement's FP64 residual of this system levels off near 1e-13
    with BEMProblem(m, Config(precision="fp32_ir", solver_tolerance=1e-12),
                    double_nodes=(ptr, idx)) as p:
        p.compute_constraints(bc)
        p.assemble_system()
        lam = p.solve_adjoint(g)
        st = p.last_stats
    err = np.abs(lam - lam_ref).max() / np.abs(lam_ref).max()
    print(f"  adjoint fp32_ir: {st['iterations']} inner it, {st['ir_sweeps']} sweeps, "
          f"rel err {err:.2e}")
    assert st["ir_sweeps"] >= 1
    assert err < 1e-8


def test_adjoint_fp32_ir_matches_fp64_10k(gpu_available):
    m = M.wigley_preset("10k")
    ptr, idx = M.double_nodes_csr(m.support_points)
    bc, _ = M.wigley_bc(m, seed=4)
    g = np.random.default_rng(13).standard_normal(m.n_dofs)
    lam = {}
    for prec in ("fp64", "fp32_ir"):
        with BEMProblem(m, Config(precision=prec), double_nodes=(ptr, idx)) as p:
            p.compute_constraints(bc)
            p.assemble_system()
            lam[prec] = p.solve_adjoint(g)
            st = p.last_stats
        print(f"  [{prec}] {st['iterations']} it, sweeps {st['ir_sweeps']}, "
              f"true rel res {st['true_rel_residual']:.2e}", flush=True)
    e = np.abs(lam["fp32_ir"] - lam["fp64"]).max() / np.abs(lam["fp64"]).max()
    print(f"  adjoint FP32-IR vs FP64 at n={m.n_dofs}: {e:.2e}")
    assert e < 1e-8


def _class_param(m, ptr, idx, nodes):
    """dofs of each chosen geometric node (its double-node class)"""
    return [idx[ptr[d]:ptr[d + 1]] for d in nodes]


def _drag_setup(m, ptr, idx, seed, n_fs, n_hull):
    """Wavy free surface (seeded elevations in [-5e-3, 5e-3] L), seeded phi_t on the hull,
    the chosen free-surface and hull nodes."""
    rng = np.random.default_rng(seed)
    fs_cells = m.cell_dofs[m.cell_material == M.FREE_SURFACE]
    fs_nodes = np.unique(fs_cells)
    hull_cells = m.cell_dofs[m.cell_material == M.HULL]
    hull_nodes = np.unique(hull_cells)
    # one representative dof per geometric node
    rep = np.array([idx[ptr[d]:ptr[d + 1]].min() for d in range(m.n_dofs)])
    fs_nodes = np.unique(rep[fs_nodes])
    hull_nodes = np.unique(rep[hull_nodes])
    fs_sel = rng.choice(fs_nodes, min(n_fs, fs_nodes.size), replace=False)
    hull_sel = rng.choice(hull_nodes, min(n_hull, hull_nodes.size), replace=False)
    X0 = m.support_points.copy()
    for d in fs_sel:
        X0[idx[ptr[d]:ptr[d + 1]], 2] += rng.uniform(-5e-3, 5e-3)
    phi_t = np.zeros(m.n_dofs)
    phi_t[np.unique(hull_cells)] = rng.uniform(-1e-3, 1e-3, np.unique(hull_cells).size)
    return X0, fs_sel, hull_sel, phi_t


def _drag_value_and_grad(m, ptr, idx, p, bc, U, X0, fs_sel, hull_sel, phi_t):
    hull = m.cell_material == M.HULL
    cells, orient = m.cell_dofs[hull], m.cell_orient[hull]
    fs_dofs = _class_param(m, ptr, idx, fs_sel)
    hull_dofs = _class_param(m, ptr, idx, hull_sel)
    th_fs = torch.zeros(len(fs_sel), dtype=torch.float64, requires_grad=True)
    th_hull = torch.zeros((len(hull_sel), 3), dtype=torch.float64, requires_grad=True)

    def support(tf, th):
        X = torch.tensor(X0, dtype=torch.float64)
        rows_f = np.concatenate(fs_dofs)
        sel_f = torch.tensor(np.repeat(np.arange(len(fs_dofs)), [len(d) for d in fs_dofs]))
        X = X.index_add(0, torch.tensor(rows_f),
                        torch.stack([torch.zeros_like(tf), torch.zeros_like(tf), tf], 1)[sel_f])
        rows_h = np.concatenate(hull_dofs)
        sel_h = torch.tensor(np.repeat(np.arange(len(hull_dofs)), [len(d) for d in hull_dofs]))
        return X.index_add(0, torch.tensor(rows_h), th[sel_h])

    def J_of(tf, th):
        X = support(tf, th)
        phi, dphi = bem_solve(X, p, bc)
        return hull_drag(X, cells, orient, phi, dphi, torch.tensor(phi_t), U)

    J = J_of(th_fs, th_hull)
    J.backward()
    g_fs, g_hull = th_fs.grad.numpy().copy(), th_hull.grad.numpy().copy()
    # the frozen-G/H form: the solution held fixed while the nodes move (the reference's
    # Jacobian re-solves with the matrices of the current geometry, free_surface.cc:4993-5014,
    # so the change of G and H with the nodes does not reach the solution)
    fr_fs = torch.zeros(len(fs_sel), dtype=torch.float64, requires_grad=True)
    fr_hull = torch.zeros((len(hull_sel), 3), dtype=torch.float64, requires_grad=True)
    with torch.no_grad():
        phi0, dphi0 = bem_solve(support(fr_fs, fr_hull), p, bc)
    X = support(fr_fs, fr_hull)
    hull_drag(X, cells, orient, phi0, dphi0, torch.tensor(phi_t), U).backward()
    frozen = (fr_fs.grad.numpy().copy(), fr_hull.grad.numpy().copy())
    return J_of, g_fs, g_hull, float(J.detach()), frozen


@pytest.mark.parametrize("preset,n_fs,n_hull,check_fs,check_hull",
                         [("2.5k", 6, 4, 6, 4), ("10k", 500, 200, 30, 12)])
def test_drag_gradient_matches_central_differences(gpu_available, preset, n_fs, n_hull,
                                                   check_fs, check_hull):
    """Gradients of the hull drag with respect to n_fs free-surface node elevations and
    n_hull hull node positions from one adjoint solve; a seeded subset (at 10k: 30
    elevations and all three components of 12 hull nodes, 66 entries) is checked against
    central differences of the GPU forward solve.  The frozen-G/H gradient (solution held
    fixed) is compared with the same differences: what it misses is the term
    -lambda^T dR/dx of the change of G and H with the nodes, whose size is reported."""
    m = M.wigley_preset(preset)
    ptr, idx = M.double_nodes_csr(m.support_points)
    bc, U = M.wigley_bc(m, seed=0)
    X0, fs_sel, hull_sel, phi_t = _drag_setup(m, ptr, idx, 7, n_fs, n_hull)
    rng = np.random.default_rng(8)
    kf = rng.choice(len(fs_sel), check_fs, replace=False)
    kh = rng.choice(len(hull_sel), check_hull, replace=False)
    with BEMProblem(m, Config(solver_tolerance=1e-13, solver_max_steps=3000),
                    double_nodes=(ptr, idx)) as p:
        p.compute_constraints(bc)
        J_of, g_fs, g_hull, J0, frozen = _drag_value_and_grad(m, ptr, idx, p, bc, U, X0,
                                                               fs_sel, hull_sel, phi_t)
        h = 1e-6
        fd_fs, fd_hull = np.zeros(check_fs), np.zeros((check_hull, 3))
        zf = torch.zeros(len(fs_sel), dtype=torch.float64)
        zh = torch.zeros((len(hull_sel), 3), dtype=torch.float64)
        with torch.no_grad():
            for j, k in enumerate(kf):
                e = zf.clone()
                e[k] = h
                fd_fs[j] = (float(J_of(e, zh)) - float(J_of(-e, zh))) / (2 * h)
            for j, k in enumerate(kh):
                for d in range(3):
                    e = zh.clone()
                    e[k, d] = h
                    fd_hull[j, d] = (float(J_of(zf, e)) - float(J_of(zf, -e))) / (2 * h)
    g = np.concatenate([g_fs[kf], g_hull[kh].ravel()])
    gf = np.concatenate([frozen[0][kf], frozen[1][kh].ravel()])
    fd = np.concatenate([fd_fs, fd_hull.ravel()])
    err = np.abs(g - fd).max() / np.abs(fd).max()
    err_frozen = np.abs(gf - fd).max() / np.abs(fd).max()
    geo = np.abs(g - gf).max() / np.abs(g).max()  # the d(G, H)/dx term, relative
    print(f"  [{preset}] n={m.n_dofs} drag {J0:.6e}; {g_fs.size} + {g_hull.size} gradient "
          f"entries; checked {fd.size}: |fd| max {np.abs(fd).max():.3e}, "
          f"adjoint vs FD max rel err {err:.2e}; frozen G/H vs FD {err_frozen:.2e} "
          f"(d(G,H)/dx term {geo:.2e} of the largest entry)")
    assert fd.size >= (60 if preset == "10k" else 18)
    assert err < 1e-6
    # the frozen form misses the geometric term, and the term is what the differences show
    assert err_frozen > 100 * err
    assert abs(np.abs(g - gf).max() - np.abs(gf - fd).max()) <= 0.05 * np.abs(g - gf).max() + \
        2 * err * np.abs(fd).max()


def test_drag_gradient_fp32_ir_matches_fp64(gpu_available):
    """The whole adjoint gradient (forward solve, transposed solve, backward of the assembly)
    in an FP32-IR context against the FP64 context's, at 10k dofs."""
    m = M.wigley_preset("10k")
    ptr, idx = M.double_nodes_csr(m.support_points)
    bc, U = M.wigley_bc(m, seed=0)
    X0, fs_sel, hull_sel, phi_t = _drag_setup(m, ptr, idx, 7, 300, 100)
    out = {}
    for prec in ("fp64", "fp32_ir"):
        with BEMProblem(m, Config(precision=prec), double_nodes=(ptr, idx)) as p:
            p.compute_constraints(bc)
            _, g_fs, g_hull, J0, _ = _drag_value_and_grad(m, ptr, idx, p, bc, U, X0, fs_sel,
                                                          hull_sel, phi_t)
        out[prec] = (np.concatenate([g_fs, g_hull.ravel()]), J0)
    g64, J64 = out["fp64"]
    g32, J32 = out["fp32_ir"]
    e = np.abs(g32 - g64).max() / np.abs(g64).max()
    print(f"  n={m.n_dofs}: {g64.size} gradient entries, FP32-IR vs FP64 {e:.2e}, "
          f"drag {abs(J32 - J64) / abs(J64):.2e}")
    assert e < 1e-8 and abs(J32 - J64) < 1e-8 * abs(J64)


@pytest.mark.parametrize("kind", ["wigley_small", "sphere4"])
def test_residual_jvp_is_transpose_of_vjp(gpu_available, kind):
    """<w, JVP(dx)> = <VJP(w), dx>: the tangent and the backward are the same derivative."""
    m, ptr, idx, H, G = T.oracle_assembly(kind)
    n = m.n_dofs
    rng = np.random.default_rng(13)
    w, phi, psi = (rng.standard_normal(n) for _ in range(3))
    dx = rng.standard_normal((n, 3))
    with BEMProblem(m, Config(), double_nodes=(ptr, idx)) as p:
        p.assemble_system()
        jv = p.residual_jvp(dx, phi, psi, free_rows_only=False)
        vj = p.residual_vjp(w, phi, psi, free_rows_only=False)
    a, b = float(w @ jv), float((vj * dx).sum())
    print(f"  [{kind}] <w, Jdx> = {a:.12e}  <J^T w, dx> = {b:.12e}")
    assert abs(a - b) <= 1e-12 * max(abs(a), abs(b))


def test_solve_rhs_matches_dense(gpu_available):
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    bc, _ = M.wigley_bc(m, seed=4)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    A, _ = T.explicit_system(H, G, m.surface_nodes, m.other_nodes, cons, bc)
    r = np.random.default_rng(14).standard_normal(m.n_dofs)
    ref = np.linalg.solve(A, r)
    with BEMProblem(m, Config(solver_tolerance=1e-13), double_nodes=(ptr, idx)) as p:
        p.compute_constraints(bc)
        p.assemble_system()
        x = p.solve_rhs(r)
    assert np.abs(x - ref).max() / np.abs(ref).max() < 1e-8


def test_solution_jvp_matches_central_differences(gpu_available):
    """Tangent of the solution along a seeded motion of free-surface and hull nodes (whole
    double-node classes) against central differences of the GPU forward solve."""
    from wavebem_amd.adjoint import bem_solve_jvp

    m = M.wigley_preset("2.5k")
    ptr, idx = M.double_nodes_csr(m.support_points)
    bc, U = M.wigley_bc(m, seed=0)
    rng = np.random.default_rng(15)
    X0 = m.support_points.copy()
    dx = np.zeros_like(X0)
    nodes = rng.choice(m.n_dofs, 60, replace=False)
    for d in nodes:
        dx[idx[ptr[d]:ptr[d + 1]]] = rng.uniform(-1.0, 1.0, 3)
    h = 1e-6
    with BEMProblem(m, Config(solver_tolerance=1e-13, solver_max_steps=3000),
                    double_nodes=(ptr, idx)) as p:
        p.compute_constraints(bc)
        phi, dphi, tphi, tdphi = bem_solve_jvp(p, X0, bc, dx)
        pp, dp, _, _ = bem_solve_jvp(p, X0 + h * dx, bc, np.zeros_like(dx))
        pm, dm, _, _ = bem_solve_jvp(p, X0 - h * dx, bc, np.zeros_like(dx))
    fd = np.concatenate([(pp - pm) / (2 * h), (dp - dm) / (2 * h)])
    jv = np.concatenate([tphi, tdphi])
    err = np.abs(jv - fd).max() / np.abs(fd).max()
    print(f"  tangent vs FD: max |fd| {np.abs(fd).max():.3e}, rel err {err:.2e}")
    assert err < 1e-6


def test_release_matrices_between_forward_and_backward(gpu_available):
    """bem_solve(release=True) frees G and H after the forward solve (the device memory
    drops by the matrix bytes, and matrix reads are refused until the next assembly); the
    backward assembles them again at the forward's geometry, so the gradient is the same,
    bit for bit, as with the matrices kept."""
    m = M.wigley_preset("2.5k")
    ptr, idx = M.double_nodes_csr(m.support_points)
    bc, U = M.wigley_bc(m, seed=0)
    hull = m.cell_material == M.HULL
    cells, orient = m.cell_dofs[hull], m.cell_orient[hull]
    phi_t = torch.zeros(m.n_dofs, dtype=torch.float64)
    n, ld = m.n_dofs, (m.n_dofs + 63) // 64 * 64
    mat_bytes = 2 * n * ld * 8                  # FP64 G and H (the adjoint is FP64)
    grads = {}
    with BEMProblem(m, Config(solver_tolerance=1e-12), double_nodes=(ptr, idx)) as p:
        p.compute_constraints(bc)
        for release in (False, True):
            X = torch.tensor(m.support_points, dtype=torch.float64, requires_grad=True)
            before = p.device_bytes()
            phi, dphi = bem_solve(X, p, bc, release=release)
            if release:
                # the first pass allocated every workspace; only the matrices go
                assert before - p.device_bytes() == mat_bytes, (before, p.device_bytes())
                with pytest.raises(WaveBemError):
                    p.compute_alpha()
            J = hull_drag(X, cells, orient, phi, dphi, phi_t, U)
            J.backward()
            grads[release] = X.grad.numpy().copy()
        p.compute_alpha()                       # the backward assembled them again
    assert np.abs(grads[False]).max() > 0
    assert np.array_equal(grads[False], grads[True])
