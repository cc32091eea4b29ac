"""Host logic of the free-surface time loop: the GMRES of the Newton systems and the fixed-step
BDF integrator, on problems with closed-form answers (no GPU)."""
import math

import numpy as np
import pytest

from wavebem_amd import timeloop as TL


def test_gmres_matches_dense_solve():
    rng = np.random.default_rng(3)
    n = 60
    A = np.eye(n) * 4.0 + rng.standard_normal((n, n)) / math.sqrt(n)
    b = rng.standard_normal(n)
    for restart in (60, 7):
        x, it, res = TL.gmres(lambda v: A @ v, b, 1e-12, max_iter=400, restart=restart)
        assert res <= 1e-12
        assert np.linalg.norm(A @ x - b) <= 1e-11
        assert np.abs(x - np.linalg.solve(A, b)).max() < 1e-10


def test_gmres_zero_rhs_and_exact_krylov_space():
    A = np.diag([1.0, 2.0, 3.0])
    x, it, res = TL.gmres(lambda v: A @ v, np.zeros(3), 1e-14)
    assert it == 0 and not x.any()
    # three distinct eigenvalues: the solution lies in the third Krylov space
    b = np.ones(3)
    x, it, res = TL.gmres(lambda v: A @ v, b, 1e-13)
    assert it <= 3 and np.allclose(x, [1.0, 0.5, 1.0 / 3.0], atol=1e-13)


class _Oscillator:
    """y' = A y, A = [[0, 1], [-w^2, 0]] (the free-surface equations of one linear wave
    mode: eta' = w-scaled phi, phi' = -g eta) as a DAE residual y' - A y."""

    def __init__(self, w):
        self.A = np.array([[0.0, 1.0], [-w * w, 0.0]])

    def residual(self, t, y, yp):
        return yp - self.A @ y

    def jacobian(self, t, y, yp, v, alpha):
        return alpha * v - self.A @ v


def _run(dt, T, w=2.0, order=2):
    dae = _Oscillator(w)
    y0 = np.array([1.0, 0.0])
    it = TL.BDFIntegrator(dae, y0, dt, max_order=order, rtol=1e-12, atol=1e-14)
    for _ in range(int(round(T / dt))):
        it.step()
    exact = np.array([math.cos(w * T), -w * math.sin(w * T)])
    return np.abs(it.y - exact).max(), it


@pytest.mark.parametrize("order,rate", [(1, 2.0), (2, 4.0)])
def test_bdf_global_order(order, rate):
    """Halving dt divides the error at T = 1 by 2 (BDF1) or 4 (BDF2); the BDF1 first step of
    the BDF2 run does not lower its order."""
    e1, _ = _run(0.01, 1.0, order=order)
    e2, _ = _run(0.005, 1.0, order=order)
    assert abs(e1 / e2 - rate) < 0.15 * rate, (e1, e2)


def test_bdf2_linear_problem_newton_in_one_iteration():
    """On a linear DAE an exact Newton step lands on the root: the second iteration's update
    is zero, so every step takes at most two iterations."""
    _, it = _run(0.01, 0.5)
    assert max(it.stats["newton_per_step"]) <= 2
    assert it.stats["steps"] == 50


def test_bdf_derivative_is_exact_for_quadratics():
    dae = _Oscillator(1.0)
    it = TL.BDFIntegrator(dae, np.array([1.0, 0.0]), 0.1)
    it.step()
    it.step()
    assert it.last_order == 2
    f = lambda t: 3.0 * t * t - 2.0 * t + 0.5  # noqa: E731
    vals = [np.array([f(0.0)]), np.array([f(0.1)]), np.array([f(0.2)])]
    assert abs(float(it.derivative(vals)[0]) - (6.0 * 0.2 - 2.0)) < 1e-12
    # the accepted step solves its DAE: y' of the BDF formula equals A y
    assert np.abs(it.last_yp - dae.A @ it.y).max() < 1e-10


def test_smooth_ramp():
    assert TL.smooth_ramp(0.0, 0.2) == 0.0
    assert TL.smooth_ramp(0.2, 0.2) == 1.0
    assert TL.smooth_ramp(5.0, 0.2) == 1.0
    assert abs(TL.smooth_ramp(0.1, 0.2) - 0.5) < 1e-15
    assert TL.smooth_ramp(0.3, 0.0) == 1.0


class _LinearStandIn(TL.FreeSurfaceStandIn):
    """The stand-in's algebra with the BEM tangent replaced by fixed matrices:
    f'_eta(v_eta, v_phi) = D v_phi + Gm v_eta (Gm only in the full form)."""

    def __init__(self, n_nodes, seed, kind):
        rng = np.random.default_rng(seed)
        self.n_nodes = n_nodes
        self.g = TL.GRAVITY
        self.jacobian_kind = kind
        Q = rng.standard_normal((n_nodes, n_nodes))
        self.D = Q @ Q.T / n_nodes + np.eye(n_nodes)   # positive, like a DtN map
        self.Gm = 0.05 * rng.standard_normal((n_nodes, n_nodes)) / math.sqrt(n_nodes)
        self.counters = dict(jvp_solves=0)

    def bem_state(self, t, y):
        return None, None, None

    def _tangent(self, phi, dphi, v_eta, v_phi, geometric):
        self.counters["jvp_solves"] += 1
        return self.D @ v_phi + (self.Gm @ v_eta if geometric else 0.0)


@pytest.mark.parametrize("kind", ["full", "frozen"])
def test_schur_newton_solve_solves_the_full_system(kind):
    dae = _LinearStandIn(40, 9, kind)
    n = dae.n_nodes
    alpha = 75.0
    J = np.zeros((2 * n, 2 * n))
    J[:n, :n] = alpha * np.eye(n) - (dae.Gm if kind == "full" else 0.0)
    J[:n, n:] = -dae.D
    J[n:, :n] = dae.g * np.eye(n)
    J[n:, n:] = alpha * np.eye(n)
    r = np.random.default_rng(10).standard_normal(2 * n)
    d, solves = dae.newton_solve(0.0, np.zeros(2 * n), None, alpha, r, 1e-12, 100)
    assert np.abs(J @ d - r).max() < 1e-11
    # the generic path on the full Jacobian agrees
    d2, it, _ = TL.gmres(lambda v: dae.jacobian(0.0, None, None, v, alpha), r, 1e-12,
                         max_iter=200)
    assert np.abs(d - d2).max() < 1e-11
    assert solves == dae.counters["jvp_solves"] - it  # one product per reduced iteration + 1
