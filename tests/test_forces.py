"""Hull force integrals and the force-file record (wavebem_amd/forces.py), pinned by closed
forms: Archimedes on a closed box (a linear pressure is exact in the Q1 space on flat
faces, so the surface integral equals the divergence-theorem volume integral), zero
moment of the hydrostatic load about the centroid, the ITTC-1957 line at a hand-computed
Reynolds number, and the record layout read back field by field."""
import math

import numpy as np
import torch

from wavebem_amd import forces as F
from wavebem_amd import mesh as M
from wavebem_amd.adjoint import hull_drag


def test_archimedes_on_a_box():
    lo, hi = np.array([-0.4, 0.1, -1.3]), np.array([0.9, 0.7, -0.2])
    m = M.box_mesh(n=5, lo=lo, hi=hi)
    g = 9.81
    p = -F.RHO * g * m.support_points[:, 2]           # hydrostatic, z up
    c = 0.5 * (lo + hi)
    force, moment, wet = F.surface_forces(m.support_points, m.cell_dofs, m.cell_orient, p, c)
    vol = np.prod(hi - lo)
    area = 2 * ((hi - lo)[0] * (hi - lo)[1] + (hi - lo)[1] * (hi - lo)[2]
                + (hi - lo)[0] * (hi - lo)[2])
    # outward normals: int p n dS = int grad p dV = -rho g V e_z
    assert np.allclose(force, [0.0, 0.0, -F.RHO * g * vol], rtol=0, atol=1e-9 * F.RHO * g * vol)
    assert np.abs(moment).max() < 1e-9 * F.RHO * g * vol
    assert abs(wet - area) < 1e-12 * area


def test_pressure_force_x_is_the_drag_functional():
    m = M.wigley_preset("2.5k")
    hull = m.cell_material == M.HULL
    cells, orient = m.cell_dofs[hull], m.cell_orient[hull]
    rng = np.random.default_rng(5)
    phi, dphi, phi_t = (rng.standard_normal(m.n_dofs) * 1e-3 for _ in range(3))
    U = 0.9
    X = torch.tensor(m.support_points)
    d = float(hull_drag(X, cells, orient, torch.tensor(phi), torch.tensor(dphi),
                        torch.tensor(phi_t), U))
    # with phi = 0 and dphi_dn = 0 the drag is the pressure -rho phi_t integrated on x
    z = torch.zeros(m.n_dofs, dtype=torch.float64)
    d0 = float(hull_drag(X, cells, orient, z, z, torch.tensor(phi_t), U))
    force, _, _ = F.surface_forces(m.support_points, cells, orient,
                                   F.linearised_pressure(phi_t, 0.0, U))
    assert abs(force[0] - d0) <= 1e-12 * max(abs(d0), 1e-30)
    assert d != d0


def test_ittc_line():
    V, L, S = 1.0, 1.0, 0.3
    re = V * L / 1.307e-6
    cd = 0.075 / (math.log10(re) - 2.0) ** 2
    f = F.ittc_viscous_force(V, L, S)
    assert abs(cd - 0.0049726) < 1e-6
    assert abs(f[0] - 0.5 * F.RHO * S * cd) < 1e-12 and f[1] == 0 and f[2] == 0


def test_force_record_layout(tmp_path):
    base = str(tmp_path / "run")
    fields = dict(press_force=(1.5, -2.25e-3, 3.0), transom_press_force=(0.0, 0.0, 0.0),
                  visc_force=(0.125, 0.0, 0.0), break_press_force=(0.0, 0.0, 0.0),
                  press_moment=(4.0, 5.0, 6.0), wet_surface=0.25,
                  press_force_test_1=(7.0, 8.0, 9.0), press_force_test_2=(1.0, 2.0, 3.123456789))
    path = F.write_force_record(base, 0.0, 0.0, **fields)
    F.write_force_record(base, 0.01, 0.0, **fields)
    F.write_force_record(base, 0.0, 0.0, **fields)          # restart truncates
    lines = open(path).read().splitlines(keepends=True)
    assert len(lines) == 1 and lines[0].endswith(" \n")
    tok = lines[0].split()
    assert len(tok) == 23
    vals = [float(t) for t in tok]
    assert vals[:4] == [0.0, 1.5, -2.25e-3, 3.0]
    assert vals[16] == 0.25
    assert tok[-1] == "3.12346"                               # 6 significant digits


def test_bernoulli_pressure_closed_forms():
    """free_surface.cc:9662-9674: a still state is hydrostatic; a stagnation point
    (grad phi = -V_inf) recovers the dynamic pressure; moving nodes see the Lagrangian
    derivative corrected by v . grad phi; the expansion in grad phi to first order is the
    linearised form -rho (phi_t + V_inf . grad phi)."""
    rng = np.random.default_rng(4)
    n = 7
    z = rng.uniform(-1.0, 0.0, n)
    V = np.array([1.3, -0.2, 0.4])
    zero = np.zeros((n, 3))
    p = F.bernoulli_pressure(np.zeros(n), zero, z, V)
    assert np.allclose(p, -F.RHO * F.GRAVITY * z, rtol=0, atol=1e-10)
    p = F.bernoulli_pressure(np.zeros(n), -np.tile(V, (n, 1)), z, V)
    assert np.allclose(p, 0.5 * F.RHO * V @ V - F.RHO * F.GRAVITY * z, rtol=0, atol=1e-10)
    gp = rng.standard_normal((n, 3))
    vel = rng.standard_normal((n, 3))
    dt = rng.standard_normal(n)
    a = F.bernoulli_pressure(dt, gp, z, V, node_vel=vel)
    b = F.bernoulli_pressure(dt - (vel * gp).sum(1), gp, z, V)
    assert np.allclose(a, b, rtol=0, atol=1e-9)
    eps = 1e-6
    lin = F.bernoulli_pressure(eps * dt, eps * gp, np.zeros(n), V) / eps
    # the second-order term is rho eps |grad phi|^2 / 2
    assert np.allclose(lin, -F.RHO * (dt + gp @ V), rtol=0, atol=1e-5 * F.RHO)
