"""Pins of the quadrature and geometry restatement (deal.II QGauss, QGaussOneOverR,
FEValues with a Q1 mapping) by textbook values, closed forms and exactness properties,
and equality of the product library's host tables with the oracle's."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cell_integrals.json")

# Abramowitz & Stegun 25.4.30, n = 4 on [-1, 1]
AS_NODES = [0.339981043584856, 0.861136311594053]
AS_WEIGHTS = [0.652145154862546, 0.347854845137454]


def test_gauss4_matches_abramowitz_stegun():
    x, w = O.qgauss_1d(4)
    assert np.all(np.diff(x) > 0)
    expect_x = sorted([0.5 * (1 - t) for t in AS_NODES] + [0.5 * (1 + t) for t in AS_NODES])
    expect_w = [0.5 * AS_WEIGHTS[1], 0.5 * AS_WEIGHTS[0], 0.5 * AS_WEIGHTS[0], 0.5 * AS_WEIGHTS[1]]
    assert np.allclose(x, expect_x, atol=1e-14)
    assert np.allclose(w, expect_w, atol=1e-14)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8])
def test_gauss_exact_to_degree_2n_minus_1(n):
    x, w = O.qgauss_1d(n)
    for k in range(2 * n):
        assert abs((w * x ** k).sum() - 1.0 / (k + 1)) < 1e-14
    # and not beyond (the degree-2n error is visible for small n)
    if n <= 4:
        k = 2 * n
        assert abs((w * x ** k).sum() - 1.0 / (k + 1)) > 1e-8


def test_gauss_2d_ordering_first_coordinate_fastest():
    p, w = O.qgauss_2d(4)
    x, _ = O.qgauss_1d(4)
    assert np.allclose(p[:4, 0], x) and np.allclose(p[:4, 1], x[0])
    assert abs(w.sum() - 1.0) < 1e-15


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("vertex", [0, 1, 2, 3])
def test_one_over_r_rule_closed_forms(vertex):
    """sum w g(p) approximates int g over the unit square with the 1/R factored out:
    int 1/R = 2 asinh(1), int x/R and int xy/R from 30-digit mpmath."""
    p, w = O.qgauss_one_over_r(5, vertex)
    assert p.shape == (50, 2) and np.all(w > 0)
    v = np.array([[0, 0], [1, 0], [0, 1], [1, 1]][vertex], float)
    r = np.linalg.norm(p - v, axis=1)
    cf = _golden()["closed_forms"]
    # local coordinates measured from the singular vertex
    s = np.abs(p[:, 0] - v[0])
    t = np.abs(p[:, 1] - v[1])
    assert abs((w / r).sum() - cf["int_inv_r_unit_square"]) < 5e-7
    assert abs((w * s / r).sum() - cf["int_x_over_r_unit_square"]) < 5e-7
    assert abs((w * s * t / r).sum() - cf["int_xy_over_r_unit_square"]) < 5e-7
    assert abs(w.sum() - 1.0) < 1e-6
    # the rule converges: order 12 reproduces the closed forms to near machine precision
    p12, w12 = O.qgauss_one_over_r(12, vertex)
    r12 = np.linalg.norm(p12 - v, axis=1)
    s12 = np.abs(p12[:, 0] - v[0])
    assert abs((w12 / r12).sum() - cf["int_inv_r_unit_square"]) < 1e-13
    assert abs((w12 * s12 / r12).sum() - cf["int_x_over_r_unit_square"]) < 1e-13
    # the singularity sits at this vertex, not at another one
    for other in range(4):
        if other == vertex:
            continue
        vo = np.array([[0, 0], [1, 0], [0, 1], [1, 1]][other], float)
        assert abs((w / np.linalg.norm(p - vo, axis=1)).sum() - cf["int_inv_r_unit_square"]) > 1e-3


def test_one_over_r_rule_vertex_rotations():
    p0, w0 = O.qgauss_one_over_r(5, 0)
    rot = {1: lambda x, y: (1 - y, x), 2: lambda x, y: (y, 1 - x), 3: lambda x, y: (1 - x, 1 - y)}
    for v, f in rot.items():
        p, w = O.qgauss_one_over_r(5, v)
        q = np.array([f(a, b) for a, b in p0])
        assert np.allclose(p, q, atol=1e-15)
        assert np.array_equal(w, w0)


def test_fe_values_flat_rectangle():
    V = np.array([[0, 0, 0], [2, 0, 0], [0, 3, 0], [2, 3, 0]], float)
    p, w = O.qgauss_2d(4)
    qp, nr, jxw, sh = O.fe_values(V, 1, p, w)
    assert abs(jxw.sum() - 6.0) < 1e-14
    assert np.allclose(nr, [0, 0, 1])
    assert np.allclose(qp[:, 0], 2 * p[:, 0]) and np.allclose(qp[:, 1], 3 * p[:, 1])
    assert np.allclose(sh.sum(axis=1), 1.0)
    _, nr2, _, _ = O.fe_values(V, -1, p, w)
    assert np.allclose(nr2, [0, 0, -1])


def test_fe_values_skewed_bilinear_area_and_moments():
    """Area and first moment of a non-planar bilinear patch against a 40-point rule."""
    V = np.array([[0, 0, 0], [1.2, 0.1, 0.0], [0.2, 0.9, 0.1], [1.1, 1.3, 0.4]])
    p, w = O.qgauss_2d(4)
    qp, nr, jxw, sh = O.fe_values(V, 1, p, w)
    P, W = O.qgauss_2d(40)
    QP, NR, JXW, _ = O.fe_values(V, 1, P, W)
    assert abs(jxw.sum() - JXW.sum()) / JXW.sum() < 1e-6
    assert np.allclose((qp * jxw[:, None]).sum(0), (QP * JXW[:, None]).sum(0), rtol=1e-6)
    assert np.allclose(np.linalg.norm(nr, axis=1), 1.0)


def test_product_quadrature_tables_equal_oracle():
    from wavebem_amd.bem import host_quadrature

    for q, s in [(4, 5), (3, 4), (6, 8)]:
        rp, rw, sp, sw = host_quadrature(q, s)
        p, w = O.qgauss_2d(q)
        assert np.array_equal(rp, p) and np.array_equal(rw, w)
        for v in range(4):
            ps, ws = O.qgauss_one_over_r(s, v)
            assert np.allclose(sp[v], ps, atol=1e-15, rtol=0)
            assert np.allclose(sw[v], ws, atol=0, rtol=1e-15)
