"""GPU runs of the free-surface time-loop stand-in (wavebem_amd.timeloop) at 2.5k dofs.

* The Newton Jacobian-vector product against central differences of the DAE right-hand
  side (each difference re-assembles G and H at the displaced mesh and solves): the full
  tangent (frozen re-solve + wb_residual_jvp / wb_solve_rhs geometric term) within 1e-6,
  and the size of the geometric term that the reference's frozen-G/H product leaves out
  (source/free_surface.cc:4993-5014).
* Five BDF steps with Newton-GMRES (the Schur-reduced Newton systems) in FP32-IR against
  FP64: state and hull force (the reference's Bernoulli pressure from the device-projected
  potential gradient) within 1e-8 of each other; and the frozen Jacobian reaching the same
  states as the full one.
"""
import functools

import numpy as np
import pytest

from wavebem_amd import mesh as M
from wavebem_amd import timeloop as TL
from wavebem_amd.bem import BEMProblem, Config

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=None)
def _mesh(name):
    m = M.wigley_preset(name)
    ptr, idx = M.double_nodes_csr(m.support_points)
    bc, U = M.wigley_bc(m, seed=0)
    return m, (ptr, idx), bc, U


def _wavy_state(dae, seed):
    """eta: a smooth seeded wave of 2e-3 L; phi: seeded values in [-1e-3, 1e-3]."""
    rng = np.random.default_rng(seed)
    eta = np.zeros(dae.n_nodes)
    phi = np.zeros(dae.n_nodes)
    x = np.bincount(dae.fs_class, weights=dae.X0[dae.fs_dofs, 0]) / dae.fs_count
    yv = np.bincount(dae.fs_class, weights=dae.X0[dae.fs_dofs, 1]) / dae.fs_count
    k1, k2, ph = rng.uniform(4.0, 8.0), rng.uniform(2.0, 6.0), rng.uniform(0, 2 * np.pi)
    eta[:] = 2e-3 * np.sin(k1 * x + ph) * np.cos(k2 * yv)
    phi[:] = rng.uniform(-1e-3, 1e-3, dae.n_nodes)
    return np.concatenate([eta, phi])


def test_jacobian_vector_product_matches_central_differences(gpu_available):
    m, dn, bc, U = _mesh("2.5k")
    with BEMProblem(m, Config(solver_tolerance=1e-13, solver_max_steps=3000),
                    double_nodes=dn) as p:
        p.compute_constraints(bc)
        dae = TL.FreeSurfaceStandIn(p, m, dn, bc)
        y = _wavy_state(dae, 5)
        t = 0.3
        rng = np.random.default_rng(6)
        v = rng.standard_normal(dae.size)
        frozen, geometric = dae.jvp_parts(t, y, v)
        jv = dae.rhs_jvp(t, y, v, kind="full")
        jv_frozen = dae.rhs_jvp(t, y, v, kind="frozen")
        h = 1e-6
        fd = (dae.rhs(t, y + h * v) - dae.rhs(t, y - h * v)) / (2 * h)
    top = slice(0, dae.n_nodes)
    scale = np.abs(fd[top]).max()
    e_full = np.abs(jv[top] - fd[top]).max() / scale
    e_frozen = np.abs(frozen - fd[top]).max() / scale
    g_rel = np.abs(geometric).max() / scale
    print(f"  n_nodes={dae.n_nodes}: J v vs central FD: full {e_full:.2e}, "
          f"frozen {e_frozen:.2e}; geometric term {g_rel:.2e} of |f' v|")
    assert e_full < 1e-6
    # one solve with both terms equals the two solves of the parts (linearity)
    assert np.abs(jv[top] - (frozen + geometric)).max() < 1e-9 * scale
    assert np.abs(jv_frozen[top] - frozen).max() < 1e-9 * scale
    # the geometric term is what the frozen product misses
    assert abs(e_frozen - g_rel) < 0.05 * g_rel + 1e-6
    assert e_frozen > 100 * e_full
    # the dynamic condition's rows are exact in both forms
    assert np.abs(jv[dae.n_nodes:] - fd[dae.n_nodes:]).max() <= 1e-7 * np.abs(
        fd[dae.n_nodes:]).max()


def _run(precision, jacobian, steps=5, dt=0.02, **newton):
    """The bench's time loop at 2.5k: the hull oscillating in surge at 1% of the Froude-0.3
    speed (period 0.5 s, ramped over 0.05 s).  The data are 100 times smaller than the
    solver's default scale, so its absolute tolerances go down by 100 as well."""
    m, dn, bc, U = _mesh("2.5k")
    bc = np.where(m.other_nodes > 0, 1e-2 * bc, bc)
    U = 1e-2 * U
    newton.setdefault("atol", 1e-11)
    with BEMProblem(m, Config(precision=precision, solver_tolerance=1e-12),
                    double_nodes=dn) as p:
        p.compute_constraints(bc)
        dae = TL.FreeSurfaceStandIn(
            p, m, dn, bc, jacobian=jacobian,
            forcing=lambda t: TL.smooth_ramp(t, 0.05) * np.cos(2 * np.pi * t / 0.5))
        it = TL.BDFIntegrator(dae, np.zeros(dae.size), dt, **newton)
        phis, forces = [np.zeros(m.n_dofs)], []
        for _ in range(steps):
            y = it.step()
            phi, _, _ = dae.bem_state(it.t, y)
            phis.append(phi.copy())
            dphi_dt = it.derivative(phis[-3:])
            eta_t = it.last_yp[:dae.n_nodes]
            forces.append(dae.hull_force(it.t, y, dphi_dt, U, eta_t)[0])
        return it.y.copy(), np.array(forces), dict(it.stats), dict(dae.counters)


def test_bdf_five_steps_fp32_ir_matches_fp64(gpu_available):
    out = {}
    for prec in ("fp64", "fp32_ir"):
        y, f, st, cnt = _run(prec, "full")
        print(f"  [{prec}] newton {st['newton_per_step']} krylov {st['krylov_per_step']} "
              f"{cnt}; force {f}")
        out[prec] = (y, f, st)
        assert max(st["newton_per_step"]) <= 4
    y64, f64, _ = out["fp64"]
    y32, f32, _ = out["fp32_ir"]
    n = y64.size // 2
    e_eta = np.abs(y32[:n] - y64[:n]).max() / np.abs(y64[:n]).max()
    e_phi = np.abs(y32[n:] - y64[n:]).max() / np.abs(y64[n:]).max()
    e_f = np.abs(f32 - f64).max() / np.abs(f64).max()
    print(f"  FP32-IR vs FP64 after 5 steps: eta {e_eta:.2e} phi {e_phi:.2e} force {e_f:.2e}")
    assert np.abs(y64[:n]).max() > 0 and np.abs(f64).max() > 0
    assert e_eta < 1e-8 and e_phi < 1e-8 and e_f < 1e-8


def test_bdf_frozen_jacobian_reaches_the_same_states(gpu_available):
    """The Jacobian only steers Newton: with the frozen form the iteration converges to
    the same root (tight Newton tolerance on both), in more iterations."""
    tight = dict(rtol=1e-10, atol=1e-13, max_newton=16)
    y_full, f_full, st_full, _ = _run("fp64", "full", steps=3, **tight)
    y_fr, f_fr, st_fr, _ = _run("fp64", "frozen", steps=3, **tight)
    e = np.abs(y_fr - y_full).max() / np.abs(y_full).max()
    print(f"  frozen vs full: state {e:.2e}; newton {st_fr['newton_per_step']} vs "
          f"{st_full['newton_per_step']}")
    assert e < 1e-8
    assert sum(st_fr["newton_per_step"]) >= sum(st_full["newton_per_step"])
