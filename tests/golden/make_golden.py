"""Writes tests/golden/cell_integrals.json: high-precision single-cell integrals.

The reference ships no known-answer data for this path, so these fixtures come from the
mathematics: the Laplace kernels of include/laplace_kernel.h:45-64 integrated against
the Q1 shape functions over one bilinear (iso-parametric) cell.  Cells away from the
collocation point: mpmath tanh-sinh at 30 digits.  Collocation point at a cell vertex: a
panelled Gauss-Legendre product rule in polar coordinates about that vertex (rho times
the integrand is smooth there), in float64, checked against a coarser panelling to 1e-12;
adaptive tanh-sinh was tried first and stops early on the thin hull cell, whose polar
integrand is sharply peaked in theta.  The oracle's Gauss 4x4 / QGaussOneOverR(5) values
must reproduce these to quadrature accuracy, and its high-order rules closely.

Entries: G_k = int phi_k / (4 pi |y-x|) dA,  H_k = int phi_k (y-x).n / (-4 pi |y-x|^3) dA,
with n = orient (a x b)/|a x b| (a = dy/dxi, b = dy/deta), k in deal.II vertex order.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import mpmath as mp

mp.mp.dps = 30
HERE = os.path.dirname(os.path.abspath(__file__))


def shape(xi, eta):
    return [(1 - xi) * (1 - eta), xi * (1 - eta), (1 - xi) * eta, xi * eta]


def geom(v, xi, eta):
    ph = shape(xi, eta)
    dxi = [-(1 - eta), (1 - eta), -eta, eta]
    deta = [-(1 - xi), -xi, (1 - xi), xi]
    y = [sum(ph[k] * v[k][d] for k in range(4)) for d in range(3)]
    a = [sum(dxi[k] * v[k][d] for k in range(4)) for d in range(3)]
    b = [sum(deta[k] * v[k][d] for k in range(4)) for d in range(3)]
    axb = [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]
    return ph, y, axb


def integrand(v, orient, x, xi, eta, k, which):
    ph, y, axb = geom(v, xi, eta)
    R = [y[d] - x[d] for d in range(3)]
    r = mp.sqrt(sum(t * t for t in R))
    jac = mp.sqrt(sum(t * t for t in axb))
    if which == "G":
        return ph[k] * jac / (4 * mp.pi * r)
    # (D . n) dA with n dA = orient * (a x b) dxi deta
    rn = orient * sum(R[d] * axb[d] for d in range(3))
    return ph[k] * rn / (-4 * mp.pi * r ** 3)


def integrate(v, orient, x, k, which):
    """Regular-cell integral by mpmath's adaptive quadrature (singular cells use
    polar_gauss_legendre)."""
    v = [[mp.mpf(t) for t in p] for p in v]
    x = [mp.mpf(t) for t in x]
    return mp.quad(lambda xi, eta: integrand(v, orient, x, xi, eta, k, which), [0, 1], [0, 1])


def polar_gauss_legendre(v, orient, sv, k, which, panels=16, npts=48):
    """Singular-cell integral in polar coordinates (rho, theta) about vertex sv of the
    reference square: rho * integrand is smooth, so a panelled Gauss-Legendre product rule
    converges fast.  (Adaptive tanh-sinh stops early here: on thin cells the integrand is
    sharply peaked in theta.)  numpy float64; convergence is checked by the caller."""
    import numpy as np

    V = np.array(v, dtype=np.float64)
    g, w = np.polynomial.legendre.leggauss(npts)
    total = 0.0
    for t0, t1, rmax in ((0.0, np.pi / 4, lambda t: 1 / np.cos(t)),
                         (np.pi / 4, np.pi / 2, lambda t: 1 / np.sin(t))):
        edges = np.linspace(t0, t1, panels + 1)
        for a, b in zip(edges[:-1], edges[1:]):
            th = 0.5 * (b - a) * (g + 1) + a
            wt = 0.5 * (b - a) * w
            for ti, wti in zip(th, wt):
                R = rmax(ti)
                # graded rho panels towards the vertex
                rb = R * np.array([0.0, 1e-3, 1e-2, 0.1, 0.4, 1.0])
                for r0, r1 in zip(rb[:-1], rb[1:]):
                    rho = 0.5 * (r1 - r0) * (g + 1) + r0
                    wr = 0.5 * (r1 - r0) * w
                    s, t = rho * np.cos(ti), rho * np.sin(ti)
                    xi = s if sv in (0, 2) else 1 - s
                    eta = t if sv in (0, 1) else 1 - t
                    ph = np.stack([(1 - xi) * (1 - eta), xi * (1 - eta), (1 - xi) * eta, xi * eta])
                    dxi = np.stack([-(1 - eta), (1 - eta), -eta, eta])
                    de = np.stack([-(1 - xi), -xi, (1 - xi), xi])
                    # y - x_sv = sum_k phi_k (v_k - v_sv): no cancellation near the vertex
                    D = V - V[sv]
                    Rv = np.einsum("kn,kd->nd", ph, D)
                    a_ = np.einsum("kn,kd->nd", dxi, V)
                    b_ = np.einsum("kn,kd->nd", de, V)
                    axb = np.cross(a_, b_)
                    r = np.linalg.norm(Rv, axis=1)
                    if which == "G":
                        val = ph[k] * np.linalg.norm(axb, axis=1) / (4 * np.pi * r)
                    else:
                        val = ph[k] * orient * (Rv * axb).sum(1) / (-4 * np.pi * r ** 3)
                    total += wti * (val * rho * wr).sum()
    return total


def cases():
    # a curved cell taken from the Wigley hull y = (B/2)(1-(2x/L)^2)(1-(z/T)^2)
    L, B, T = 1.0, 0.1, 0.0625

    def hull(x, z):
        return [x, 0.5 * B * (1 - (2 * x / L) ** 2) * (1 - (z / T) ** 2), z]

    x0, x1, z0, z1 = 0.10, 0.18, -0.030, -0.018
    hull_cell = [hull(x0, z0), hull(x1, z0), hull(x0, z1), hull(x1, z1)]
    # a skewed flat-ish quad on the free surface
    fs_cell = [[0.55, 0.02, 0.0], [0.63, 0.025, 0.0], [0.56, 0.09, 0.0], [0.65, 0.10, 0.001]]
    out = []
    out.append(dict(name="hull_far", vertices=hull_cell, orient=-1,
                    x=[0.30, 0.035, -0.02], singular_vertex=None))
    out.append(dict(name="hull_near", vertices=hull_cell, orient=-1,
                    x=[0.20, 0.045, -0.024], singular_vertex=None))
    out.append(dict(name="fs_far", vertices=fs_cell, orient=1,
                    x=[0.40, 0.30, 0.0], singular_vertex=None))
    for sv in range(4):
        out.append(dict(name=f"hull_singular_v{sv}", vertices=hull_cell, orient=-1,
                        x=hull_cell[sv], singular_vertex=sv))
    out.append(dict(name="fs_singular_v1", vertices=fs_cell, orient=1, x=fs_cell[1],
                    singular_vertex=1))
    return out


def main():
    res = []
    for c in cases():
        vals = {}
        sv = c["singular_vertex"]
        for which in ("G", "H"):
            if sv is None:
                vals[which] = [float(integrate(c["vertices"], c["orient"], c["x"], k, which))
                               for k in range(4)]
                c["method"] = "mpmath tanh-sinh, 30 digits"
            else:
                fine = [polar_gauss_legendre(c["vertices"], c["orient"], sv, k, which, 16, 48)
                        for k in range(4)]
                coarse = [polar_gauss_legendre(c["vertices"], c["orient"], sv, k, which, 8, 32)
                          for k in range(4)]
                scale = max(abs(t) for t in fine)
                assert max(abs(a - b) for a, b in zip(fine, coarse)) < 1e-12 * scale, (which, c)
                vals[which] = fine
                c["method"] = "polar Gauss-Legendre (16 panels x 48, graded rho), float64"
        c = dict(c)
        c.update(vals)
        print(c["name"], vals, file=sys.stderr, flush=True)
        res.append(c)
    closed = {
        # int_{[0,1]^2} dA / sqrt(x^2+y^2) = 2 asinh(1)
        "int_inv_r_unit_square": float(2 * mp.asinh(1)),
        # int x / r dA over the unit square (mpmath, polar split)
        "int_x_over_r_unit_square": float(mp.quad(lambda x, y: x / mp.sqrt(x * x + y * y),
                                                  [0, 1], [0, 1])),
        "int_xy_over_r_unit_square": float(mp.quad(lambda x, y: x * y / mp.sqrt(x * x + y * y),
                                                   [0, 1], [0, 1])),
    }
    with open(os.path.join(HERE, "cell_integrals.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py (mpmath, 30 digits)",
                   "cases": res, "closed_forms": closed}, f, indent=1)


if __name__ == "__main__":
    main()
