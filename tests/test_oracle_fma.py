"""CPU restatement of BEMFMA (oracle/bem_fma.c), pinned by mathematics.

* Normalised associated Legendre functions sqrt((n-m)!/(n+m)!) P_n^m against
  scipy.special.lpmv, and their sin-weighted derivative against central differences.
* Every expansion operator against the exact potentials it approximates: a unit charge
  (1/|x-y|) and a unit normal dipole ((x-y).n/|x-y|^3), through P2M -> M2P, P2M -> M2M ->
  M2P, P2M -> M2L -> L2P and P2M -> M2L -> L2L -> L2P; the error falls with the truncation
  order p as the separation ratio^(p+1).
* The octree lists: each (collocation node, quadrature point) pair is used exactly once,
  by the near field (a singular cell covering all its points), M2L -> L2L -> L2P or M2P.
* The near field is the dense assembly: with one octree level everything is near and the
  products equal H phi and G dphi_dn of the oracle's assemble_system to rounding.
* The products converge to the dense ones as p grows.
"""
import numpy as np
import pytest
from scipy.special import factorial, lpmv

import oracle as O
from tests import helpers as T


@pytest.mark.parametrize("n,m,x", [(0, 0, 0.3), (1, 1, 0.3), (2, 1, -0.4), (5, 3, 0.77),
                                   (10, 7, 0.1), (14, 0, -0.93), (20, 20, 0.5)])
def test_legendre(n, m, x):
    v, ds = O.fma_legendre(n, m, x)
    ref = lpmv(m, n, x) * np.sqrt(factorial(n - m) / factorial(n + m))
    assert abs(v - ref) <= 1e-13 * max(1.0, abs(ref))
    h = 1e-6
    fd = (O.fma_legendre(n, m, x + h)[0] - O.fma_legendre(n, m, x - h)[0]) / (2 * h)
    assert abs(ds - fd * np.sqrt(1 - x * x)) <= 1e-7 * max(1.0, abs(ds))


_SRC = np.array([0.1, -0.05, 0.08])
_X = np.array([1.5, 0.7, -0.9])
_NRM = np.array([0.3, 0.4, 0.866]) / np.linalg.norm([0.3, 0.4, 0.866])


@pytest.mark.parametrize("kind,centres", [
    (0, [[0, 0, 0]]),                                         # P2M -> M2P
    (1, [[0, 0, 0]]),                                         # dipole P2M -> M2P
    (2, [[0, 0, 0], [0.3, 0.2, 0.1]]),                        # M2M
    (3, [[0, 0, 0], [1.3, 0.6, -0.8]]),                       # M2L -> L2P
    (4, [[0, 0, 0], [1.3, 0.6, -0.8], [1.4, 0.65, -0.85]]),   # M2L -> L2L -> L2P
])
def test_expansions_converge_to_exact_potential(kind, centres):
    R = _X - _SRC
    exact = R @ _NRM / np.linalg.norm(R) ** 3 if kind == 1 else 1.0 / np.linalg.norm(R)
    errs = [abs(O.fma_expansion_check(p, kind, _SRC, centres, _X, normal=_NRM) - exact)
            / abs(exact) for p in (2, 4, 6, 8, 10)]
    assert all(b < a for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-8, errs


@pytest.mark.parametrize("kind", ["box6", "sphere4", "wigley_small"])
@pytest.mark.parametrize("levels", [3, 6, 10])
def test_every_pair_used_exactly_once(kind, levels):
    m, ptr, idx, _, _ = T.oracle_assembly(kind)
    f = O.FMA(m.support_points, m.cell_dofs, m.cell_orient, ptr, idx, trunc_order=2,
              num_levels=levels)
    try:
        c = f.coverage(m.n_cells)
    finally:
        f.close()
    assert c.min() == 1 and c.max() == 1


@pytest.mark.parametrize("kind", ["box6", "sphere4", "wigley_small"])
def test_single_level_is_the_dense_product(kind):
    m, ptr, idx, H, G = T.oracle_assembly(kind)
    rng = np.random.default_rng(31)
    phi, psi = rng.standard_normal(m.n_dofs), rng.standard_normal(m.n_dofs)
    f = O.FMA(m.support_points, m.cell_dofs, m.cell_orient, ptr, idx, trunc_order=4,
              num_levels=1)
    pn, pd = f.products(phi, psi)
    f.close()
    assert np.abs(pn - H @ phi).max() <= 1e-13 * np.abs(H @ phi).max()
    assert np.abs(pd - G @ psi).max() <= 1e-13 * np.abs(G @ psi).max()


def test_products_converge_with_truncation_order():
    m, ptr, idx, H, G = T.oracle_assembly("wigley_small")
    rng = np.random.default_rng(32)
    phi, psi = rng.standard_normal(m.n_dofs), rng.standard_normal(m.n_dofs)
    hp, gp = H @ phi, G @ psi
    eh, eg = [], []
    for p in (2, 4, 8):
        f = O.FMA(m.support_points, m.cell_dofs, m.cell_orient, ptr, idx, trunc_order=p,
                  num_levels=10)
        pn, pd = f.products(phi, psi)
        f.close()
        eh.append(np.abs(pn - hp).max() / np.abs(hp).max())
        eg.append(np.abs(pd - gp).max() / np.abs(gp).max())
    assert eh[0] > 4 * eh[1] > 16 * eh[2] and eg[0] > 4 * eg[1] > 16 * eg[2], (eh, eg)
    assert eh[2] < 1e-3 and eg[2] < 1e-4
