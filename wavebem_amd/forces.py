"""Hull force integrals and the force-file record of FreeSurface::compute_pressure.

The reference integrates, over the wall (hull) cells and the 4x4 Gauss points of each,
  press_force  = sum_q p(q) n_q JxW_q          (source/free_surface.cc:9697-9725)
  wet_surface  = sum_q JxW_q
  press_moment = sum_q p(q) (x_q - b) x n_q JxW_q   (:9599-9604, b the hull barycentre)
with p(q) the nodal pressure interpolated by the Q1 shape functions and n_q the cell
normal, adds the ITTC-1957 viscous drag (:10121-10126) and appends one line per time to
"<output_file_name>_force.txt" (:10128-10139).  ``bernoulli_pressure`` is the reference's
nodal pressure (:9662-9674) from the complete potential gradient
(BEMProblem.compute_potential_gradients, on the device) and the time derivative of phi;
``linearised_pressure`` is the single-state form of the adjoint configuration,
p = -rho (phi_t + U dphi/dx), as used by adjoint.hull_drag.
"""
from __future__ import annotations

import math

import numpy as np

RHO = 1025.1          # source/free_surface.cc:9417
GRAVITY = 9.81        # source/free_surface.cc:9416 (g of compute_pressure)
NU_WATER = 1.307e-6   # kinematic viscosity at 10 C, source/free_surface.cc:10122


def _rule(q):
    t, w = np.polynomial.legendre.leggauss(q)
    x = 0.5 * (t + 1.0)
    w = 0.5 * w
    xi = np.tile(x, q)            # point i + q*j: xi fastest (deal.II QGauss<2> order)
    eta = np.repeat(x, q)
    return xi, eta, np.repeat(w, q) * np.tile(w, q)


def _cell_quadrature(support, cells, orient, q_order):
    """Points (k, nq, 3), unit normals (k, nq, 3), JxW (k, nq), shape values (nq, 4)."""
    xi, eta, w = _rule(q_order)
    sh = np.stack([(1 - xi) * (1 - eta), xi * (1 - eta), (1 - xi) * eta, xi * eta], 1)
    dx = np.stack([-(1 - eta), (1 - eta), -eta, eta], 1)
    de = np.stack([-(1 - xi), -xi, (1 - xi), xi], 1)
    V = np.asarray(support, dtype=np.float64)[np.asarray(cells)]      # (k, 4, 3)
    X = np.einsum("qv,cvd->cqd", sh, V)
    a = np.einsum("qv,cvd->cqd", dx, V)
    b = np.einsum("qv,cvd->cqd", de, V)
    cr = np.cross(a, b)
    cl = np.linalg.norm(cr, axis=2)
    o = np.asarray(orient, dtype=np.float64)
    n = o[:, None, None] * cr / cl[:, :, None]
    return X, n, cl * w[None, :], sh


def surface_forces(support, cells, orient, p_nodes, barycentre=(0.0, 0.0, 0.0), q_order=4):
    """(press_force (3,), press_moment (3,), wet_surface) of the nodal pressure ``p_nodes``
    over ``cells`` (k x 4 dofs, deal.II vertex order, orient = +-1)."""
    X, n, jxw, sh = _cell_quadrature(support, cells, orient, q_order)
    p = np.asarray(p_nodes, dtype=np.float64)[np.asarray(cells)] @ sh.T     # (k, nq)
    f = (p * jxw)[:, :, None] * n
    force = f.sum(axis=(0, 1))
    arm = X - np.asarray(barycentre, dtype=np.float64)[None, None, :]
    moment = np.cross(arm, f).sum(axis=(0, 1))
    return force, moment, float(jxw.sum())


def bernoulli_pressure(dphi_dt, grad_phi, z, V_inf, node_vel=None, rho=RHO, g=GRAVITY):
    """Nodal pressure of FreeSurface::compute_pressure (source/free_surface.cc:9662-9674):
        p = 1/2 rho |V_inf|^2 - 1/2 rho |V_inf + grad phi|^2 - rho g z
            - rho (DphiDt - v . grad phi),
    with grad phi the complete potential gradient at the nodes (n, 3;
    compute_potential_gradients), DphiDt the time derivative of phi following the node,
    v the node velocity (n, 3; None = fixed nodes, DphiDt = phi_t) and z the nodes'
    vertical coordinate."""
    gp = np.asarray(grad_phi, dtype=np.float64)
    V = np.asarray(V_inf, dtype=np.float64)
    tot = V[None, :] + gp
    p = (0.5 * rho * float(V @ V) - 0.5 * rho * (tot * tot).sum(1)
         - rho * g * np.asarray(z, dtype=np.float64))
    conv = 0.0 if node_vel is None else (np.asarray(node_vel, dtype=np.float64) * gp).sum(1)
    return p - rho * (np.asarray(dphi_dt, dtype=np.float64) - conv)


def linearised_pressure(phi_t, dphi_dx, U, rho=RHO):
    """p = -rho (phi_t + U dphi/dx) at the nodes (the adjoint configuration's form)."""
    return -rho * (np.asarray(phi_t) + U * np.asarray(dphi_dx))


def ittc_viscous_force(V, L, wet_surface, rho=RHO, nu=NU_WATER):
    """ITTC-1957 friction line: Re = max(1e-5, V L / nu), Cd = 0.075 / (log10 Re - 2)^2,
    drag 0.5 rho V^2 S Cd along x (source/free_surface.cc:10121-10126)."""
    re = max(1e-5, V * L / nu)
    cd = 0.075 / (math.log10(re) - 2.0) ** 2
    return np.array([0.5 * rho * V * V * wet_surface * cd, 0.0, 0.0])


def _g(v):
    # operator<< of a double on a default std::ostream: 6 significant digits, %g style
    return f"{float(v):g}"


def force_line(t, press_force, transom_press_force, visc_force, break_press_force,
               press_moment, wet_surface, press_force_test_1, press_force_test_2):
    """One record of the force file, fields as written by the reference:
    t, press_force, transom_press_force, visc_force, break_press_force, press_moment,
    wet surface (hull + transom), press_force_test_1, press_force_test_2, then " \\n"."""
    parts = [_g(t)]
    for v in (press_force, transom_press_force, visc_force, break_press_force, press_moment):
        parts.append(" ".join(_g(c) for c in v))
    parts.append(_g(wet_surface))
    for v in (press_force_test_1, press_force_test_2):
        parts.append(" ".join(_g(c) for c in v))
    return " ".join(parts) + " \n"


def write_force_record(output_file_name, t, initial_time, **fields):
    """Appends the record to "<output_file_name>_force.txt"; the file is started afresh when
    |t - initial_time| < 1e-5, as in the reference."""
    path = output_file_name + "_force.txt"
    mode = "w" if abs(t - initial_time) < 1e-5 else "a"
    with open(path, mode) as f:
        f.write(force_line(t, **fields))
    return path
