"""GPU parity of the entry points and configurations the other GPU tests leave out:

* wb_solve_system_dev (device pointers, caller order; the step bench.py times) against a
  dense LU solve of the oracle's constrained system at 2.5k dofs, and against
  wb_solve_system on the same context at 41k dofs;
* the 20k-dof Wigley configuration in FP64: whole rows against the oracle, and FP32-IR
  against FP64 on the same system;
* BEMProblem::reinit on a live context (source/bem_problem.cc:55-71, called again after
  every remesh): 2.5k -> 10k -> 2.5k dofs on one context, each solve against a fresh
  context and against dense LU;
* wb_solve (BEMProblem::solve, source/bem_problem.cc:969-987) against assemble_system +
  solve_system.

Bars: entries 1e-10 (row-scaled), solutions 1e-8 relative to the largest value.
"""
import functools

import numpy as np
import pytest

import oracle as O
from tests import helpers as T
from wavebem_amd import mesh as M
from wavebem_amd.bem import BEMProblem, Config

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=None)
def _mesh(name):
    m = M.wigley_preset(name)
    ptr, idx = M.double_nodes_csr(m.support_points)
    return m, ptr, idx


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _dense_reference(kind, bc):
    m, ptr, idx, H, G = T.oracle_assembly(kind)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    return T.dense_solution(H, G, m.surface_nodes, m.other_nodes, cons, bc)


def _solve_dev(p, m, bc):
    import torch

    bc_d = torch.tensor(bc, dtype=torch.float64, device="cuda")
    phi_d = torch.tensor(bc * m.surface_nodes, dtype=torch.float64, device="cuda")
    dphi_d = torch.tensor(bc * m.other_nodes, dtype=torch.float64, device="cuda")
    st = p.solve_system_dev(phi_d.data_ptr(), dphi_d.data_ptr(), bc_d.data_ptr())
    torch.cuda.synchronize()
    return phi_d.cpu().numpy(), dphi_d.cpu().numpy(), st


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_solve_system_dev_matches_dense(gpu_available, precision):
    m, ptr, idx, H, G = T.oracle_assembly("wigley_2.5k")
    bc, _ = M.wigley_bc(m, seed=4)
    x = _dense_reference("wigley_2.5k", bc)
    print(f"[solve_system_dev] 2.5k n={m.n_dofs} {precision}", flush=True)
    with BEMProblem(m, Config(precision=precision, solver_tolerance=1e-12),
                    double_nodes=(ptr, idx)) as p:
        p.compute_constraints(bc)
        p.assemble_system()
        phi, dphi, st = _solve_dev(p, m, bc)
    sol = np.where(m.other_nodes > 0, phi, dphi)
    e = _rel(sol, x)
    # the known entries are left as given
    S = m.surface_nodes > 0
    assert np.array_equal(phi[S], bc[S]) and np.array_equal(dphi[~S], bc[~S])
    print(f"  {st['iterations']} it, sweeps {st['ir_sweeps']}: solution vs dense LU {e:.2e}")
    assert e < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_solve_system_dev_matches_host_path_40k(gpu_available, precision):
    """The device-pointer solve (the bench step) and the host-buffer solve run the same
    solver on one context: their solutions agree to the last bits the solver leaves."""
    m, ptr, idx = _mesh("40k")
    bc, _ = M.wigley_bc(m, seed=0)
    print(f"[solve_system_dev] 40k n={m.n_dofs} {precision}", flush=True)
    with BEMProblem(m, Config(precision=precision), double_nodes=(ptr, idx)) as p:
        p.compute_constraints(bc)
        p.assemble_system()
        phi_h, dphi_h = bc * m.surface_nodes, bc * m.other_nodes
        st_h = p.solve_system(phi_h, dphi_h, bc)
        phi_d, dphi_d, st_d = _solve_dev(p, m, bc)
    assert st_d["iterations"] == st_h["iterations"]
    e_phi, e_dn = _rel(phi_d, phi_h), _rel(dphi_d, dphi_h)
    print(f"  {st_d['iterations']} it: phi {e_phi:.2e} dphi/dn {e_dn:.2e}")
    assert e_phi < 1e-12 and e_dn < 1e-12


def test_rows_match_oracle_20k(gpu_available):
    """The 20k configuration, FP64: 40 whole rows (5 blocks of 8 spread over the device
    numbering) against the oracle's row-subset assembly, and alpha on those rows."""
    m, ptr, idx = _mesh("20k")
    n = m.n_dofs
    starts = [0, n // 3, n // 2, (2 * n) // 3, n - 8]
    blocks = [(s, s + 8) for s in starts]
    rows = np.concatenate([np.arange(a, b) for a, b in blocks]).astype(np.int32)
    print(f"[20k rows] n={n} cells={m.n_cells}", flush=True)
    with BEMProblem(m, Config(), double_nodes=(ptr, idx)) as p:
        p.assemble_system()
        Gd, Hd = p.matrix_rows(rows)
        alpha = p.compute_alpha()
    worst, r = 0.0, 0
    for a, b in blocks:
        Ho, Go = O.assemble_rows(m.support_points, m.cell_dofs, m.cell_orient, ptr, idx, a, b)
        for k in range(b - a):
            for D, R in ((Gd[r + k], Go[k]), (Hd[r + k], Ho[k])):
                worst = max(worst, float(np.abs(D - R).max() / np.abs(R).max()))
            assert abs(alpha[a + k] + Ho[k].sum()) < 1e-11
        r += b - a
    print(f"  worst row-scaled entry error {worst:.2e} over {rows.size} rows")
    assert worst < 1e-10


def test_fp32_ir_matches_fp64_20k(gpu_available):
    m, ptr, idx = _mesh("20k")
    bc, _ = M.wigley_bc(m, seed=1)
    out = {}
    for prec in ("fp64", "fp32_ir"):
        phi, dphi = bc * m.surface_nodes, bc * m.other_nodes
        with BEMProblem(m, Config(precision=prec), double_nodes=(ptr, idx)) as p:
            p.compute_constraints(bc)
            p.assemble_system()
            st = p.solve_system(phi, dphi, bc)
        print(f"  [20k {prec}] {st['iterations']} it, sweeps {st['ir_sweeps']}, "
              f"true rel res {st['true_rel_residual']:.2e}", flush=True)
        out[prec] = (phi, dphi)
    N = m.other_nodes > 0
    e_phi = _rel(out["fp32_ir"][0][N], out["fp64"][0][N])
    e_dn = _rel(out["fp32_ir"][1][~N], out["fp64"][1][~N])
    print(f"  FP32-IR vs FP64: phi {e_phi:.2e} dphi/dn {e_dn:.2e}")
    assert e_phi < 1e-8 and e_dn < 1e-8


def _solve_on(p, m, bc):
    phi, dphi = bc * m.surface_nodes, bc * m.other_nodes
    p.compute_constraints(bc)
    p.assemble_system()
    st = p.solve_system(phi, dphi, bc)
    return np.where(m.other_nodes > 0, phi, dphi), st


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_reinit_on_live_context(gpu_available, precision):
    """One context through 2.5k -> 10k -> 2.5k dofs (buffers grow, then the mesh shrinks
    back): every solve equals a fresh context's, bit for bit, and the 2.5k solves match
    dense LU of the oracle's system."""
    cfg = dict(precision=precision, solver_tolerance=1e-12)
    seq = ["2.5k", "10k", "2.5k"]
    ref = {}
    for name in ("2.5k", "10k"):
        m, ptr, idx = _mesh(name)
        bc, _ = M.wigley_bc(m, seed=2)
        with BEMProblem(m, Config(**cfg), double_nodes=(ptr, idx)) as p:
            ref[name] = _solve_on(p, m, bc)[0]
    x = _dense_reference("wigley_2.5k", M.wigley_bc(_mesh("2.5k")[0], seed=2)[0])
    with BEMProblem(config=Config(**cfg)) as p:
        for name in seq:
            m, ptr, idx = _mesh(name)
            bc, _ = M.wigley_bc(m, seed=2)
            print(f"[reinit] {precision} -> {name} n={m.n_dofs}", flush=True)
            p.reinit(m, (ptr, idx))
            sol, st = _solve_on(p, m, bc)
            assert np.array_equal(sol, ref[name]), (name, _rel(sol, ref[name]))
            if name == "2.5k":
                e = _rel(sol, x)
                print(f"  {st['iterations']} it: vs dense LU {e:.2e}")
                assert e < 1e-8


def test_wb_solve_matches_assemble_and_solve(gpu_available):
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    bc, _ = M.wigley_bc(m, seed=3)
    with BEMProblem(m, Config(solver_tolerance=1e-12), double_nodes=(ptr, idx)) as p:
        p.compute_constraints(bc)
        sol_a, _ = _solve_on(p, m, bc)
        phi, dphi = bc * m.surface_nodes, bc * m.other_nodes
        st = p.solve(phi, dphi, bc)
    sol_b = np.where(m.other_nodes > 0, phi, dphi)
    assert np.array_equal(sol_a, sol_b)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    x = T.dense_solution(H, G, m.surface_nodes, m.other_nodes, cons, bc)
    assert _rel(sol_b, x) < 1e-8
    assert st["iterations"] > 0


@pytest.mark.parametrize("precision,inner,bar", [("fp64", 1, 1e-12), ("fp32_ir", 0, 1e-8),
                                                 ("fp32_ir", 1, 5e-5), ("fp32_ir", 2, 1e-7)])
def test_coarse_operator_is_PtAP(gpu_available, precision, inner, bar):
    """The coarse operator of the preconditioner against P^T A P formed from the oracle's
    matrices: A P from the FP64 matrices (FP64), from the FP32 hi part (FP32-IR, FP32 inner
    operator), from the aggregate sums the FP16 copy's build writes (FP32-IR, FP16 inner
    operator; FP16 rounding of the entries, hence the looser bar) and from the FP32 sums the
    compressed operator's sketch pass writes (FP32-IR, compressed inner operator).  Measured:
    9e-15, 5e-10, 7e-6 and 2e-9.  A wrong aggregate map or
    reduction would only cost iterations in a solve; here it fails."""
    from wavebem_amd import bem as B

    m, ptr, idx, H, G = T.oracle_assembly("wigley_2.5k")
    bc, _ = M.wigley_bc(m, seed=5)
    with BEMProblem(m, Config(precision=precision, ir_inner_format=inner),
                    double_nodes=(ptr, idx)) as p:
        _solve_on(p, m, bc)
        Ac, agg = p.coarse_matrix()
    n = m.n_dofs
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    S, N = m.surface_nodes, m.other_nodes
    A = H * N[None, :] - G * S[None, :]
    A[np.arange(n), np.arange(n)] += -H.sum(axis=1) * N
    for r in cons.rows:
        A[r, :] = 0.0
        A[r, r] = 1.0
    a_of = B.host_strip_plan(m)["perm"] // agg  # aggregate of each dof
    nc = Ac.shape[0]
    assert nc == (n + agg - 1) // agg and a_of.max() == nc - 1
    P = np.zeros((n, nc))
    P[np.arange(n), a_of] = 1.0
    ref = P.T @ A @ P
    e = float(np.abs(Ac - ref).max() / np.abs(ref).max())
    print(f"  [{precision} inner {inner}] nc={nc} agg={agg}: |Ac - P^T A P| / max {e:.2e}")
    assert e < bar
