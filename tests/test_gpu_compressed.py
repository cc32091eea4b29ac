"""The compressed inner operator of FP32 iterative refinement (ir_inner_format 2,
wb_hmat.hip) on the GPU.

* its product against the FP64-accurate operator (hi + lo) of the same context, next to the
  FP32 hi operator and the row-scaled FP16 copy: the compressed copy is a correction-solve
  operator, so it is held to the FP16 copy's order of accuracy (a bar of 1e-4, five times
  its measured error), not to FP64's;
* solutions: the refinement residual always uses the FP64-accurate operator, so the
  solutions meet the same bars as every FP32-IR solve: 1e-8 against dense LU of the
  oracle's system (2.5k dofs) and against the FP64 solve (10k dofs).
"""
import numpy as np
import pytest

import oracle as O
from tests import helpers as T
from wavebem_amd import mesh as M
from wavebem_amd.bem import BEMProblem, Config

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["wigley_2.5k", "wigley_10k"])
def test_inner_operator_error(gpu_available, kind):
    m = T.wigley_preset(kind.split("_")[1])
    ptr, idx = M.double_nodes_csr(m.support_points)
    bc, _ = M.wigley_bc(m, seed=6)
    x = np.random.default_rng(6).standard_normal(m.n_dofs)
    errs = {}
    # measured: 3.5e-8 (FP32 hi), 1.2e-5 - 1.4e-5 (FP16 copy), 1.7e-5 - 1.9e-5 (compressed,
    # rank threshold 1e-4 of each block's largest entry)
    for fmt, bar in ((0, 2e-7), (1, 5e-5), (2, 1e-4)):
        with BEMProblem(m, Config(precision="fp32_ir", ir_inner_format=fmt),
                        double_nodes=(ptr, idx)) as p:
            p.compute_constraints(bc)
            p.assemble_system()
            ref = p.constrained_vmult(x)
            y = p.inner_vmult(x)
        errs[fmt] = float(np.linalg.norm(y - ref) / np.linalg.norm(ref))
        print(f"  [{kind}] inner format {fmt}: |A_inner x - A x| / |A x| = {errs[fmt]:.2e}",
              flush=True)
        assert errs[fmt] < bar


def test_compressed_solution_matches_dense(gpu_available):
    m, ptr, idx, H, G = T.oracle_assembly("wigley_2.5k")
    bc, _ = M.wigley_bc(m, seed=7)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    x = T.dense_solution(H, G, m.surface_nodes, m.other_nodes, cons, bc)
    phi, dphi = bc * m.surface_nodes, bc * m.other_nodes
    with BEMProblem(m, Config(precision="fp32_ir", ir_inner_format=2, solver_tolerance=1e-12),
                    double_nodes=(ptr, idx)) as p:
        p.compute_constraints(bc)
        p.assemble_system()
        st = p.solve_system(phi, dphi, bc)
    sol = np.where(m.other_nodes > 0, phi, dphi)
    e = float(np.abs(sol - x).max() / np.abs(x).max())
    print(f"  {st['iterations']} it, sweeps {st['ir_sweeps']} "
          f"{st['sweep_iterations'][:st['ir_sweeps']]}: vs dense LU {e:.2e}")
    assert e < 1e-8


def test_compressed_matches_fp64_10k(gpu_available):
    m = T.wigley_preset("10k")
    ptr, idx = M.double_nodes_csr(m.support_points)
    bc, _ = M.wigley_bc(m, seed=8)
    out = {}
    for name, kw in (("fp64", dict(precision="fp64")),
                     ("compressed", dict(precision="fp32_ir", ir_inner_format=2))):
        phi, dphi = bc * m.surface_nodes, bc * m.other_nodes
        with BEMProblem(m, Config(**kw), double_nodes=(ptr, idx)) as p:
            p.compute_constraints(bc)
            p.assemble_system()
            st = p.solve_system(phi, dphi, bc)
        print(f"  [{name}] {st['iterations']} it, sweeps {st['ir_sweeps']}", flush=True)
        out[name] = (phi, dphi)
    N = m.other_nodes > 0
    e_phi = np.abs(out["compressed"][0][N] - out["fp64"][0][N]).max() / np.abs(out["fp64"][0][N]).max()
    e_dn = np.abs(out["compressed"][1][~N] - out["fp64"][1][~N]).max() / np.abs(out["fp64"][1][~N]).max()
    print(f"  compressed vs FP64: phi {e_phi:.2e} dphi/dn {e_dn:.2e}")
    assert e_phi < 1e-8 and e_dn < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_short_restarts_match_dense(gpu_available, precision):
    """GMRES restarted every 4 vectors (gmres_max_tmp_vectors = 6): the restart residuals
    and the first product of every cycle take the inputs that are not fresh Krylov vectors
    (cluster-order gather, separate coarse restriction), the steps queued ahead of the host
    end cycles; the solution must still meet the dense-LU bar (SolverGMRES restarts,
    source/bem_problem.cc:821-895)."""
    m, ptr, idx, H, G = T.oracle_assembly("wigley_2.5k")
    bc, _ = M.wigley_bc(m, seed=9)
    cons = O.compute_constraints(ptr, idx, m.surface_nodes, bc)
    x = T.dense_solution(H, G, m.surface_nodes, m.other_nodes, cons, bc)
    phi, dphi = bc * m.surface_nodes, bc * m.other_nodes
    kw = dict(precision=precision, solver_tolerance=1e-12, gmres_max_tmp_vectors=6,
              solver_max_steps=2000)
    if precision == "fp32_ir":
        kw["ir_inner_format"] = 2
    with BEMProblem(m, Config(**kw), double_nodes=(ptr, idx)) as p:
        p.compute_constraints(bc)
        p.assemble_system()
        st = p.solve_system(phi, dphi, bc)
    sol = np.where(m.other_nodes > 0, phi, dphi)
    e = float(np.abs(sol - x).max() / np.abs(x).max())
    print(f"  [{precision}] {st['iterations']} it (restart length 4): vs dense LU {e:.2e}",
          flush=True)
    assert st["iterations"] > 8
    assert e < 1e-8
