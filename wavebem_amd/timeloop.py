"""Free-surface time loop around the BEM solve: a stand-in for WaveBEM's DAE integration.

The reference integrates its free-surface DAE with SUNDIALS IDA (variable-order BDF, Newton
iterations, SPGMR on a user Jacobian-vector product: source/dae_time_integrator.cc:54-110,
373-386).  Its residual (FreeSurface::residual, source/free_surface.cc:4815-4900) moves the
free-surface nodes, re-assembles G and H and solves the BEM system at every evaluation; its
Jacobian-vector product (FreeSurface::jacobian, source/free_surface.cc:4915-5053) re-solves
the BEM system with the matrices of the current geometry held fixed ("frozen G/H") and adds
sparse derivative matrices of the other equations.

This module keeps that structure for a deliberately simple free-surface model, so that the
GPU path runs inside a real Newton loop with the same call pattern:

* state y = (eta, phi): the vertical displacement eta_c of every free-surface node (each
  class of coincident dofs moves together) and its potential phi_c;
* right-hand side f(t, y) = (w, -g eta), the kinematic and dynamic conditions of linear wave
  theory at zero forward speed, w_c = dphi/dn at the node from the BEM solve on the
  displaced mesh with Dirichlet data phi on the free surface and the Neumann data of the
  hull times a factor of time (``forcing(t) * bc``: a ramp, or a ramped oscillation that
  keeps the waves small against the cells, which linear theory assumes);
* DAE residual F(t, y, y') = y' - f(t, y) (FreeSurface::residual's role) and Jacobian
  alpha v - f'(t, y) v (FreeSurface::jacobian's role, alpha = dF/dy');
* f' v in two forms: ``"full"``, the exact tangent, whose free-surface rows come from the
  frozen re-solve with data v_phi plus the geometric term -A_c^-1 (dR/dx dx) of moving the
  nodes by v_eta (wb_residual_jvp, wb_solve_rhs); and ``"frozen"``, the reference's form,
  without the geometric term.

What is not modelled: the nonlinear (semi-Lagrangian) free-surface conditions, forward speed
in the free-surface conditions, the hull's rigid-body equations, remeshing, and IDA's
variable order, variable step and error control (the integrator is BDF of fixed step, order
1 for the first step and 2 after).  The BEM part (assembly, constraints, GMRES in FP64 or
FP32 iterative refinement) is the production path.
"""
from __future__ import annotations

import math

import numpy as np

from . import mesh as M

GRAVITY = 9.81  # source/free_surface.cc:9332


def smooth_ramp(t, t_ramp):
    """0 at t = 0 rising to 1 at t_ramp with zero slope at both ends."""
    if t_ramp <= 0.0 or t >= t_ramp:
        return 1.0
    return 0.5 * (1.0 - math.cos(math.pi * max(t, 0.0) / t_ramp))


class FreeSurfaceStandIn:
    """The DAE of the free-surface stand-in on one BEMProblem context (FP64 or FP32-IR).

    problem      a BEMProblem built on ``mesh`` with the same double nodes
    double_nodes (ptr, idx) CSR of coincident dofs
    bc           boundary data: Neumann entries used as given times forcing(t), Dirichlet
                 entries ignored (they come from phi)
    jacobian     "full" or "frozen"
    forcing      time factor of the Neumann data; default the smooth ramp to 1 at t_ramp
    """

    def __init__(self, problem, mesh, double_nodes, bc, t_ramp=0.2, g=GRAVITY,
                 jacobian="full", forcing=None):
        if jacobian not in ("full", "frozen"):
            raise ValueError("jacobian must be 'full' or 'frozen'")
        self.p = problem
        self.mesh = mesh
        self.g = float(g)
        self.t_ramp = float(t_ramp)
        # time factor of the Neumann data (and of the stream of the hull pressure)
        self.forcing = forcing if forcing is not None else (
            lambda t: smooth_ramp(t, self.t_ramp))
        self.jacobian_kind = jacobian
        ptr, idx = double_nodes
        n = mesh.n_dofs
        self.n = n
        fs = np.flatnonzero(mesh.dof_material == M.FREE_SURFACE)
        if fs.size == 0:
            raise ValueError("mesh has no free-surface dofs")
        rep = np.array([idx[ptr[d]] for d in range(n)], dtype=np.int64)  # smallest dof of class
        reps, cls_of_fs = np.unique(rep[fs], return_inverse=True)
        self.n_nodes = reps.size
        self.fs_dofs = fs
        self.fs_class = cls_of_fs
        self.fs_count = np.bincount(cls_of_fs, minlength=self.n_nodes).astype(np.float64)
        # every dof of a free-surface class moves with it
        rows, cls = [], []
        for c, r in enumerate(reps):
            members = idx[ptr[r]:ptr[r + 1]]
            rows.append(members)
            cls.append(np.full(members.size, c))
        self.move_rows = np.concatenate(rows)
        self.move_class = np.concatenate(cls)
        self.X0 = mesh.support_points.copy()
        self.bc = np.asarray(bc, dtype=np.float64).copy()
        self.dirichlet = problem.surface_nodes > 0
        if not np.all(self.dirichlet[fs]):
            raise ValueError("free-surface dofs must be Dirichlet dofs")
        self._cache_key = None
        self._cache = None
        self.counters = dict(assemblies=0, solves=0, jvp_solves=0)

    # ---- state <-> mesh and boundary data
    @property
    def size(self):
        return 2 * self.n_nodes

    def split(self, y):
        return y[:self.n_nodes], y[self.n_nodes:]

    def support(self, eta):
        X = self.X0.copy()
        X[self.move_rows, 2] += eta[self.move_class]
        return X

    def boundary_data(self, t, phi_nodes):
        tmp = self.bc * self.forcing(t)
        tmp[self.dirichlet] = 0.0
        tmp[self.fs_dofs] = phi_nodes[self.fs_class]
        return tmp

    def node_mean(self, v):
        """Mean of a dof vector over the free-surface dofs of each node."""
        return np.bincount(self.fs_class, weights=v[self.fs_dofs],
                           minlength=self.n_nodes) / self.fs_count

    # ---- the BEM state of (t, y): assembly at the displaced mesh and one solve
    def bem_state(self, t, y):
        key = (float(t), y.tobytes())
        if self._cache_key == key:
            return self._cache
        eta, phi_n = self.split(y)
        tmp = self.boundary_data(t, phi_n)
        self.p.set_support_points(self.support(eta))
        self.p.assemble_system()
        phi = tmp * self.p.surface_nodes
        dphi = tmp * self.p.other_nodes
        self.p.solve_system(phi, dphi, tmp)
        self.counters["assemblies"] += 1
        self.counters["solves"] += 1
        self._cache_key = key
        self._cache = (phi, dphi, tmp)
        return self._cache

    def rhs(self, t, y):
        phi, dphi, _ = self.bem_state(t, y)
        eta, _ = self.split(y)
        return np.concatenate([self.node_mean(dphi), -self.g * eta])

    # ---- DAE interface (FreeSurface::residual / FreeSurface::jacobian roles)
    def residual(self, t, y, yp):
        return yp - self.rhs(t, y)

    def _tangent(self, phi, dphi, v_eta, v_phi, geometric):
        """dphi/dn at the free-surface nodes of the BEM tangent for a change v_phi of the
        Dirichlet data and, with ``geometric``, a vertical move v_eta of the nodes, in one
        solve of the constrained system:
            A_c du = b(v_phi) - dR/dx dx,
        b(v_phi) the right-hand side (constraint inhomogeneities included) of the data
        v_phi on the free surface and zero Neumann data (the negative of wb_residual at
        zero unknowns), dR/dx dx the tangent of the BEM rows (wb_residual_jvp).  Without
        ``geometric`` this is the reference's product, a re-solve with G and H frozen."""
        b = np.zeros(self.n)
        if np.any(v_phi):
            dphi_data = np.zeros(self.n)
            dphi_data[self.fs_dofs] = v_phi[self.fs_class]
            b -= self.p.residual(dphi_data, np.zeros(self.n))
        if geometric and np.any(v_eta):
            dx = np.zeros((self.n, 3))
            dx[self.move_rows, 2] = v_eta[self.move_class]
            b -= self.p.residual_jvp(dx, phi, dphi, free_rows_only=True)
        if not b.any():
            return np.zeros(self.n_nodes)
        du = self.p.solve_rhs(b)
        self.counters["jvp_solves"] += 1
        # du holds dphi/dn on the Dirichlet (free-surface) dofs
        return self.node_mean(du)

    def jvp_parts(self, t, y, v):
        """(frozen, geometric): the free-surface rows of f'(t, y) v split into the re-solve
        with frozen G/H and data v_phi, and the term of moving the nodes by v_eta (two
        solves; the product itself takes one)."""
        phi, dphi, _ = self.bem_state(t, y)
        v_eta, v_phi = self.split(v)
        zero = np.zeros(self.n_nodes)
        return (self._tangent(phi, dphi, v_eta, v_phi, False),
                self._tangent(phi, dphi, v_eta, zero, True))

    def rhs_jvp(self, t, y, v, kind=None):
        """f'(t, y) v, with the geometric term ("full") or without it ("frozen")."""
        kind = kind or self.jacobian_kind
        phi, dphi, _ = self.bem_state(t, y)
        v_eta, v_phi = self.split(v)
        top = self._tangent(phi, dphi, v_eta, v_phi, kind == "full")
        return np.concatenate([top, -self.g * v_eta])

    def newton_solve(self, t, y, yp, alpha, r, tol, max_iter):
        """Solves (alpha I - f') d = r by eliminating the dynamic rows: they give
        d_phi = (r_phi - g d_eta) / alpha, and d_eta solves the Schur complement
            S d_eta = alpha d_eta - f'_eta(d_eta, -(g / alpha) d_eta)
                    = r_eta + f'_eta(0, r_phi / alpha)
        by GMRES (one BEM solve per product, half the unknowns of the full system, and a
        spectrum alpha + g k / alpha near the positive reals).  The residual of the full
        system equals that of the reduced one.  Returns (d, BEM solves)."""
        phi, dphi, _ = self.bem_state(t, y)
        geo = self.jacobian_kind == "full"
        r_eta, r_phi = self.split(r)
        c = self.g / alpha
        zero = np.zeros(self.n_nodes)
        rhs = r_eta + self._tangent(phi, dphi, zero, r_phi / alpha, False)
        S = lambda u: alpha * u - self._tangent(phi, dphi, u, -c * u, geo)  # noqa: E731
        d_eta, k, _ = gmres(S, rhs, tol, max_iter=max_iter)
        d_phi = (r_phi - self.g * d_eta) / alpha
        return np.concatenate([d_eta, d_phi]), k + 1

    def jacobian(self, t, y, yp, v, alpha):
        """(dF/dy + alpha dF/dy') v = alpha v - f'(t, y) v."""
        return alpha * v - self.rhs_jvp(t, y, v)

    # ---- hull force of the current BEM state
    def hull_force(self, t, y, dphi_dt, U, eta_t=None):
        """Pressure force on the hull (FreeSurface::compute_pressure's press_force,
        source/free_surface.cc:9662-9725): the nodal Bernoulli pressure of
        forces.bernoulli_pressure with the complete potential gradient projected on the
        device (wb_compute_potential_gradients), V_inf = (U forcing(t), 0, 0) (the stream
        whose Neumann data the hull carries), DphiDt = dphi_dt following the nodes and node
        velocity (0, 0, eta_t) on the moving nodes, integrated over the hull cells at the
        displaced mesh.  Returns the force (3,)."""
        from .forces import bernoulli_pressure, surface_forces

        phi, dphi, _ = self.bem_state(t, y)
        eta, _ = self.split(y)
        X = self.support(eta)
        grad = self.p.compute_potential_gradients(phi, dphi)
        vel = None
        if eta_t is not None:
            vel = np.zeros((self.n, 3))
            vel[self.move_rows, 2] = eta_t[self.move_class]
        V = np.array([U * self.forcing(t), 0.0, 0.0])
        p_nodes = bernoulli_pressure(dphi_dt, grad, X[:, 2], V, node_vel=vel)
        hull = self.mesh.cell_material == M.HULL
        force, _, _ = surface_forces(X, self.mesh.cell_dofs[hull], self.mesh.cell_orient[hull],
                                     p_nodes)
        return force

    def hull_force_x(self, t, y, phi_t, U):
        """x-force of the linearised Bernoulli pressure -rho (phi_t + U forcing(t) dphi/dx)
        over the hull at the displaced mesh (adjoint.hull_drag)."""
        import torch

        from .adjoint import hull_drag

        phi, dphi, _ = self.bem_state(t, y)
        eta, _ = self.split(y)
        X = self.support(eta)
        hull = self.mesh.cell_material == M.HULL
        return float(hull_drag(torch.from_numpy(X), self.mesh.cell_dofs[hull],
                               self.mesh.cell_orient[hull], torch.from_numpy(phi),
                               torch.from_numpy(dphi), torch.from_numpy(phi_t),
                               U * self.forcing(t)))


def gmres(apply, b, tol, max_iter=40, restart=40):
    """Restarted GMRES (modified Gram-Schmidt, Givens rotations) from x = 0 for
    ||b - A x|| <= tol; returns (x, iterations, residual norm)."""
    n = b.size
    x = np.zeros(n)
    beta = float(np.linalg.norm(b))
    it = 0
    if beta <= tol:
        return x, 0, beta
    r = b.copy()
    while it < max_iter:
        m = min(restart, max_iter - it)
        V = np.zeros((m + 1, n))
        Hh = np.zeros((m + 1, m))
        cs, sn = np.zeros(m), np.zeros(m)
        gam = np.zeros(m + 1)
        gam[0] = beta
        V[0] = r / beta
        k_done = 0
        for k in range(m):
            w = apply(V[k])
            it += 1
            for j in range(k + 1):
                Hh[j, k] = float(V[j] @ w)
                w = w - Hh[j, k] * V[j]
            Hh[k + 1, k] = float(np.linalg.norm(w))
            if Hh[k + 1, k] > 0.0:
                V[k + 1] = w / Hh[k + 1, k]
            for j in range(k):
                a, c = Hh[j, k], Hh[j + 1, k]
                Hh[j, k], Hh[j + 1, k] = cs[j] * a + sn[j] * c, -sn[j] * a + cs[j] * c
            d = math.hypot(Hh[k, k], Hh[k + 1, k])
            cs[k], sn[k] = (Hh[k, k] / d, Hh[k + 1, k] / d) if d > 0 else (1.0, 0.0)
            Hh[k, k] = d
            Hh[k + 1, k] = 0.0
            gam[k + 1] = -sn[k] * gam[k]
            gam[k] = cs[k] * gam[k]
            k_done = k + 1
            if abs(gam[k + 1]) <= tol:
                break
        yk = np.linalg.solve(np.triu(Hh[:k_done, :k_done]), gam[:k_done])
        x = x + V[:k_done].T @ yk
        beta = abs(gam[k_done])
        if beta <= tol:
            break
        r = b - apply(x)
        it += 1
        beta = float(np.linalg.norm(r))
    return x, it, beta


class BDFIntegrator:
    """Fixed-step BDF (order 1 for the first step, then 2) with Newton iterations whose
    linear systems are solved by GMRES on the DAE's Jacobian-vector product, the structure
    of IDA with SPGMR and a user J v (source/dae_time_integrator.cc:54-110, 373-386).

    Newton stops when the weighted RMS norm of the update, weights 1 / (rtol |y| + atol),
    is below ``newton_tol`` (IDA's convergence test has the same form; atol sits above the
    noise that the BEM solver's absolute tolerance leaves in the residual); GMRES solves each
    Newton system to ``krylov_rtol`` times the norm of the residual, on the DAE's own
    ``newton_solve`` when it has one (and ``structured``), else on ``jacobian``."""

    COEF = {1: (1.0, (-1.0,)), 2: (1.5, (-2.0, 0.5))}

    def __init__(self, dae, y0, dt, t0=0.0, max_order=2, rtol=1e-6, atol=1e-9,
                 newton_tol=1.0, max_newton=6, krylov_rtol=1e-6, max_krylov=40,
                 structured=True):
        self.dae = dae
        self.dt = float(dt)
        self.t = float(t0)
        self.max_order = int(max_order)
        self.hist = [np.asarray(y0, dtype=np.float64).copy()]
        self.rtol, self.atol = rtol, atol
        self.newton_tol, self.max_newton = newton_tol, max_newton
        self.krylov_rtol, self.max_krylov = krylov_rtol, max_krylov
        self.structured = structured
        self.stats = dict(steps=0, newton=0, krylov=0, newton_per_step=[], krylov_per_step=[])

    @property
    def y(self):
        return self.hist[-1]

    def restart(self, y, t):
        """Continues from state y at time t (a restored dump) with an empty history: the
        next step is of order 1, as IDA's restart from (y, y')."""
        self.hist = [np.asarray(y, dtype=np.float64).copy()]
        self.t = float(t)

    def yp_of(self, y):
        """y' of the BDF formula for y at t + dt, and alpha = dy'/dy."""
        q = min(self.max_order, len(self.hist))
        a0, a = self.COEF[q]
        c = sum(aj * self.hist[-1 - j] for j, aj in enumerate(a))
        return (a0 * y + c) / self.dt, a0 / self.dt

    def step(self):
        t1 = self.t + self.dt
        q = min(self.max_order, len(self.hist))
        y = self.hist[-1].copy() if q == 1 else 2.0 * self.hist[-1] - self.hist[-2]
        w = 1.0 / (self.rtol * np.abs(self.hist[-1]) + self.atol)
        nk = nn = 0
        converged = False
        trace = []
        for _ in range(self.max_newton):
            yp, alpha = self.yp_of(y)
            F = self.dae.residual(t1, y, yp)
            fn = float(np.linalg.norm(F))
            if self.structured and hasattr(self.dae, "newton_solve"):
                d, k = self.dae.newton_solve(t1, y, yp, alpha, -F, self.krylov_rtol * fn,
                                             self.max_krylov)
            else:
                d, k, _ = gmres(lambda v: self.dae.jacobian(t1, y, yp, v, alpha), -F,
                                self.krylov_rtol * fn, max_iter=self.max_krylov)
            y = y + d
            nk += k
            nn += 1
            dw = math.sqrt(float(np.mean((d * w) ** 2)))
            trace.append((fn, dw, k))
            if dw <= self.newton_tol:
                converged = True
                break
        if not converged:
            raise RuntimeError(f"Newton did not converge at t = {t1}")
        self.last_order = q
        self.last_yp, _ = self.yp_of(y)
        self.hist.append(y)
        if len(self.hist) > self.max_order:
            self.hist.pop(0)
        self.t = t1
        self.stats["steps"] += 1
        self.stats["newton"] += nn
        self.stats["krylov"] += nk
        self.stats["newton_per_step"].append(nn)
        self.stats["krylov_per_step"].append(nk)
        self.last_trace = trace  # per Newton iteration: |F|, WRMS norm of the update, GMRES
        return y

    def derivative(self, values):
        """Time derivative at the last step of a quantity given at its last steps
        (``values[-1]`` at t, ``values[-2]`` at t - dt, ...), by the BDF formula of the
        order that step used."""
        q = min(getattr(self, "last_order", 1), len(values) - 1)
        if q < 1:
            return np.zeros_like(values[-1])
        a0, a = self.COEF[q]
        return (a0 * values[-1] + sum(aj * values[-2 - j] for j, aj in enumerate(a))) / self.dt
