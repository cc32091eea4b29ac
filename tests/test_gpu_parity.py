"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Bars (from the hot-path contract): G and H entries within 1e-10 of the oracle relative to
the row scale, alpha within 1e-12, operator products within 1e-12, solutions (phi and
dphi/dn) within 1e-8 relative.
"""
import numpy as np
import pytest

import oracle as O
from tests import helpers as T
from wavebem_amd import mesh as M
from wavebem_amd.bem import BEMProblem, Config, NoConvergence

pytestmark = pytest.mark.gpu


def _problem(m, ptr, idx, **cfg):
    p = BEMProblem(m, Config(**cfg), double_nodes=(ptr, idx))
    p.assemble_system()
    return p


@pytest.mark.parametrize("kind", ["wigley_small", "box6", "sphere4"])
def test_assembly_matches_oracle(gpu_available, kind):
    m, ptr, idx, H, G = T.oracle_assembly(kind)
    print(f"[assembly parity] {kind} n={m.n_dofs}", flush=True)
    with _problem(m, ptr, idx) as p:
        Gd, Hd = p.matrices()
        alpha = p.compute_alpha()
    eg = T.row_scaled_err(Gd, G)
    eh = T.row_scaled_err(Hd, H)
    ea = np.abs(alpha - O.compute_alpha(H)).max()
    print(f"  G {eg:.2e} H {eh:.2e} alpha {ea:.2e}")
    assert eg < 1e-10
    assert eh < 1e-10
    assert ea < 1e-12


@pytest.mark.parametrize("kind,strips,precision", [
    ("mixed", 1, "fp64"), ("mixed", 1, "fp32_ir"), ("tri_block", 1, "fp64"),
    ("wigley_small", 0, "fp64"), ("wigley_small", 0, "fp32_ir")])
def test_assembly_strip_and_tile_kernels(gpu_available, kind, strips, precision):
    """The strip kernel (structured patches), the column-tile kernel (the rest) and both
    in one launch sequence, against the oracle."""
    m, ptr, idx, H, G = T.oracle_assembly(kind)
    print(f"[assembly kernels] {kind} strips={strips} {precision} n={m.n_dofs}", flush=True)
    bar = 1e-10 if precision == "fp64" else 1e-9
    with _problem(m, ptr, idx, assembly_strips=strips, precision=precision) as p:
        Gd, Hd = p.matrices()
        alpha = p.compute_alpha()
    eg = T.row_scaled_err(Gd, G)
    eh = T.row_scaled_err(Hd, H)
    ea = np.abs(alpha - O.compute_alpha(H)).max()
    print(f"  G {eg:.2e} H {eh:.2e} alpha {ea:.2e}")
    assert eg < bar
    assert eh < bar
    assert ea < 1e-12


def test_assembly_fp32_ir_storage(gpu_available):
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    with _problem(m, ptr, idx, precision="fp32_ir") as p:
        Gd, Hd = p.matrices()
        alpha = p.compute_alpha()
    # hi (FP32) + lo (bf16) keeps about 2^-33 relative per entry
    assert T.row_scaled_err(Gd, G) < 1e-9
    assert T.row_scaled_err(Hd, H) < 1e-9
    assert np.abs(alpha - O.compute_alpha(H)).max() < 1e-12


def test_vmult_rhs_match_oracle(gpu_available):
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    rng = np.random.default_rng(3)
    src = rng.standard_normal(m.n_dofs)
    with _problem(m, ptr, idx) as p:
        alpha = O.compute_alpha(H)
        for name, fd, fo in (("vmult", p.vmult, O.vmult), ("rhs", p.compute_rhs, O.compute_rhs)):
            d = fd(src)
            o = fo(G, H, alpha, m.surface_nodes, m.other_nodes, src)
            err = np.abs(d - o).max() / np.abs(o).max()
            print(f"  {name} rel err {err:.2e}")
            assert err < 1e-12


def test_constrained_vmult_matches_oracle(gpu_available):
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    bc, _ = M.wigley_bc(m)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    rng = np.random.default_rng(4)
    src = rng.standard_normal(m.n_dofs)
    with _problem(m, ptr, idx) as p:
        nl = p.compute_constraints(bc)
        assert nl == cons.n_lines
        d = p.constrained_vmult(src)
    o = O.constrained_vmult(G, H, O.compute_alpha(H), m.surface_nodes, m.other_nodes, cons, src)
    assert np.abs(d - o).max() / np.abs(o).max() < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir", "fp32_ir_inner32"])
def test_solve_system_matches_dense_oracle(gpu_available, precision):
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    bc, _ = M.wigley_bc(m, seed=0)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    x = T.dense_solution(H, G, m.surface_nodes, m.other_nodes, cons, bc)
    phi = bc * m.surface_nodes
    dphi = bc * m.other_nodes
    # fp32_ir: inner solves on the FP16 operator (the default); fp32_ir_inner32: on FP32
    kw = dict(precision="fp32_ir", ir_inner_format=0) if precision == "fp32_ir_inner32" \
        else dict(precision=precision)
    with _problem(m, ptr, idx, solver_tolerance=1e-13, **kw) as p:
        p.compute_constraints(bc)
        st = p.solve_system(phi, dphi, bc)
    print(f"  [{precision}] stats {st}")
    N = m.other_nodes > 0
    S = ~N
    e_phi = np.abs(phi[N] - x[N]).max() / np.abs(x[N]).max()
    e_dn = np.abs(dphi[S] - x[S]).max() / np.abs(x[S]).max()
    print(f"  phi err {e_phi:.2e} dphi_dn err {e_dn:.2e}")
    assert e_phi < 1e-8
    assert e_dn < 1e-8
    # known entries untouched
    assert np.array_equal(phi[S], bc[S])
    assert np.array_equal(dphi[N], bc[N])


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_coarse_space_keeps_the_solution(gpu_available, precision):
    """The coarse-space correction of the block-Jacobi preconditioner (on by default) and
    plain block Jacobi (precond_coarse = 0) reach the dense solution on the 2.5k tank, the
    former in fewer GMRES iterations."""
    m = M.wigley_preset("2.5k")
    ptr, idx = M.double_nodes_csr(m.support_points)
    H, G = O.assemble(m.support_points, m.cell_dofs, m.cell_orient, ptr, idx, n_threads=8)
    bc, _ = M.wigley_bc(m, seed=4)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    x = T.dense_solution(H, G, m.surface_nodes, m.other_nodes, cons, bc)
    its = {}
    for coarse in (0, 128):
        phi = bc * m.surface_nodes
        dphi = bc * m.other_nodes
        with _problem(m, ptr, idx, precision=precision, solver_tolerance=1e-12,
                      precond_coarse=coarse) as p:
            p.compute_constraints(bc)
            st = p.solve_system(phi, dphi, bc)
        sol = np.where(m.surface_nodes == 1, dphi, phi)
        err = np.abs(sol - x).max() / np.abs(x).max()
        print(f"  [{precision} coarse {coarse}] {st['iterations']} it, err {err:.2e}")
        assert err < 1e-8
        its[coarse] = st["iterations"]
    assert its[128] < its[0]


@pytest.mark.parametrize("block", [32, 100, 128])
def test_coarse_space_with_other_block_sizes(gpu_available, block):
    """Block-Jacobi blocks of 32, 100 and 128 rows with the coarse space (a block of up to
    128 rows meets up to three 64-dof aggregates; 128-row blocks run two wavefronts):
    the FP32-IR solve reaches the dense solution of the 2.5k tank."""
    m = M.wigley_preset("2.5k")
    ptr, idx = M.double_nodes_csr(m.support_points)
    H, G = O.assemble(m.support_points, m.cell_dofs, m.cell_orient, ptr, idx, n_threads=8)
    bc, _ = M.wigley_bc(m, seed=6)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    x = T.dense_solution(H, G, m.surface_nodes, m.other_nodes, cons, bc)
    phi = bc * m.surface_nodes
    dphi = bc * m.other_nodes
    with _problem(m, ptr, idx, precision="fp32_ir", solver_tolerance=1e-12,
                  precond_block=block) as p:
        p.compute_constraints(bc)
        st = p.solve_system(phi, dphi, bc)
    sol = np.where(m.surface_nodes == 1, dphi, phi)
    err = np.abs(sol - x).max() / np.abs(x).max()
    print(f"  [block {block}] {st['iterations']} it, err {err:.2e}")
    assert err < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_coarse_space_192_at_10k(gpu_available, precision):
    """More than 128 coarse unknowns (the coarse inverse eliminated in FP32 in LDS): at
    10,951 dofs the 192-unknown coarse space and plain block Jacobi give the same solution
    to 1e-9 (relative), the former in fewer iterations."""
    m = M.wigley_preset("10k")
    ptr, idx = M.double_nodes_csr(m.support_points)
    bc, _ = M.wigley_bc(m, seed=5)
    out, its = {}, {}
    for coarse in (0, 192):
        phi = bc * m.surface_nodes
        dphi = bc * m.other_nodes
        with _problem(m, ptr, idx, precision=precision, solver_tolerance=1e-12,
                      precond_coarse=coarse) as p:
            p.compute_constraints(bc)
            st = p.solve_system(phi, dphi, bc)
        out[coarse] = np.where(m.surface_nodes == 1, dphi, phi)
        its[coarse] = st["iterations"]
    err = np.abs(out[192] - out[0]).max() / np.abs(out[0]).max()
    print(f"  [{precision}] iterations {its}, solutions differ by {err:.2e}")
    assert m.n_dofs > 128 * 64 and its[192] < its[0]
    assert err < 1e-9


def test_fp16_inner_hands_over_to_fp32(gpu_available):
    """An FP16 sweep that does not halve the outer residual (forced here by an inner
    reduction of 0.9 per sweep) switches the remaining sweeps to the FP32 operator; the
    solve still reaches the dense solution."""
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    bc, _ = M.wigley_bc(m, seed=2)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    x = T.dense_solution(H, G, m.surface_nodes, m.other_nodes, cons, bc)
    phi = bc * m.surface_nodes
    dphi = bc * m.other_nodes
    with _problem(m, ptr, idx, precision="fp32_ir", ir_inner_format=1,
                  ir_inner16_reduction=0.9, solver_tolerance=1e-13) as p:
        p.compute_constraints(bc)
        st = p.solve_system(phi, dphi, bc)
    print(f"  sweeps {st['ir_sweeps']} {st['sweep_iterations'][:st['ir_sweeps']]}")
    sol = np.where(m.surface_nodes == 1, dphi, phi)
    assert np.abs(sol - x).max() / np.abs(x).max() < 1e-8


def test_solve_matches_oracle_gmres(gpu_available):
    """Same converged solution as the oracle's reference-preconditioned GMRES."""
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    bc, _ = M.wigley_bc(m, seed=1)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    r = O.solve_system(G, H, m.surface_nodes, m.other_nodes, cons, bc, phi=bc * m.surface_nodes,
                       dphi_dn=bc * m.other_nodes, tol=1e-14, precond=1)
    assert r.status == 0
    phi = bc * m.surface_nodes
    dphi = bc * m.other_nodes
    with _problem(m, ptr, idx, solver_tolerance=1e-14) as p:
        p.compute_constraints(bc)
        p.solve_system(phi, dphi, bc)
    assert np.abs(phi - r.phi).max() / np.abs(r.phi).max() < 1e-8
    assert np.abs(dphi - r.dphi_dn).max() / np.abs(r.dphi_dn).max() < 1e-8


def test_residual_matches_oracle(gpu_available):
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    bc, _ = M.wigley_bc(m, seed=2)
    rng = np.random.default_rng(5)
    phi = rng.standard_normal(m.n_dofs)
    dphi = rng.standard_normal(m.n_dofs)
    tmp = dphi * m.other_nodes + phi * m.surface_nodes
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, tmp)
    o = O.residual(G, H, m.surface_nodes, m.other_nodes, cons, phi, dphi)
    with _problem(m, ptr, idx) as p:
        p.compute_constraints(tmp)
        d = p.residual(phi, dphi)
    assert np.abs(d - o).max() / np.abs(o).max() < 1e-12


def test_no_convergence_raises_and_leaves_outputs(gpu_available):
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    bc, _ = M.wigley_bc(m)
    phi = bc * m.surface_nodes
    dphi = bc * m.other_nodes
    phi0, dphi0 = phi.copy(), dphi.copy()
    with _problem(m, ptr, idx, solver_tolerance=0.0, solver_max_steps=2) as p:
        p.compute_constraints(bc)
        with pytest.raises(NoConvergence):
            p.solve_system(phi, dphi, bc)
    assert np.array_equal(phi, phi0) and np.array_equal(dphi, dphi0)


def test_pure_neumann_vmult(gpu_available):
    """All-Neumann masks trigger dst.add(-||dst||) (source/bem_problem.cc:667-668)."""
    m, ptr, idx, H, G = T.oracle_assembly("box6")
    surf = np.zeros(m.n_dofs)
    oth = np.ones(m.n_dofs)
    rng = np.random.default_rng(6)
    src = rng.standard_normal(m.n_dofs)
    with _problem(m, ptr, idx) as p:
        p.set_masks(surf, oth)
        d = p.vmult(src)
    o = O.vmult(G, H, O.compute_alpha(H), surf, oth, src)
    assert np.abs(d - o).max() / np.abs(o).max() < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir", "fp32_ir_fp16"])
@pytest.mark.parametrize("k", [5, 17])
def test_solve_system_multi_matches_dense_oracle(gpu_available, precision, k):
    """k seeded right-hand sides solved together (per-column GMRES state, batches of at most
    16; the FP32-IR inner solves on the compressed operator, read once per group of 8
    vectors, or with fp32_ir_fp16 on the FP16 copy's batched matrix-core product); every
    column against the dense solution of its own constrained system, to the same 1e-8 bar
    as a single solve."""
    kw = dict(precision=precision)
    if precision == "fp32_ir_fp16":
        kw = dict(precision="fp32_ir", ir_inner_format=1)
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    bcs = [M.wigley_bc(m, seed=s)[0] * (1.0 + 0.25 * s) for s in range(k)]
    tmp = np.stack(bcs)
    phi = tmp * m.surface_nodes
    dphi = tmp * m.other_nodes
    with _problem(m, ptr, idx, solver_tolerance=1e-13, **kw) as p:
        p.compute_constraints(bcs[0])
        st = p.solve_system_multi(phi, dphi, tmp)
    print(f"  [{precision} k={k}] stats {st}")
    # the reported worst true residual (FP32-IR: the refinement's last r = b - A x with the
    # FP64-accurate operator; FP64: a fresh product) is a converged solve's
    assert 0.0 < st["true_rel_residual"] < 1e-10, st
    N = m.other_nodes > 0
    S = ~N
    for v in range(k):
        cons = O.compute_constraints(ptr, idx, m.surface_nodes, bcs[v])
        x = T.dense_solution(H, G, m.surface_nodes, m.other_nodes, cons, bcs[v])
        e_phi = np.abs(phi[v, N] - x[N]).max() / np.abs(x[N]).max()
        e_dn = np.abs(dphi[v, S] - x[S]).max() / np.abs(x[S]).max()
        assert e_phi < 1e-8, (v, e_phi)
        assert e_dn < 1e-8, (v, e_dn)
        assert np.array_equal(phi[v, S], bcs[v][S])
        assert np.array_equal(dphi[v, N], bcs[v][N])


def test_solve_system_multi_rejects_pure_neumann(gpu_available):
    m, ptr, idx, H, G = T.oracle_assembly("box6")
    with _problem(m, ptr, idx) as p:
        p.set_masks(np.zeros(m.n_dofs), np.ones(m.n_dofs))
        z = np.zeros((2, m.n_dofs))
        with pytest.raises(Exception, match="Neumann"):
            p.solve_system_multi(z.copy(), z.copy(), z)
