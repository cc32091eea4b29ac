"""Shared test helpers: small meshes, oracle runs, explicit constrained systems."""
from __future__ import annotations

import functools

import numpy as np

import oracle as O
from wavebem_amd import mesh as M

SMALL_WIGLEY = dict(nh=12, nt=4, nu=6, nd=8, ny=8, nzd=3)


@functools.lru_cache(maxsize=None)
def wigley_small():
    return M.wigley_tank(**SMALL_WIGLEY)


@functools.lru_cache(maxsize=None)
def wigley_preset(name):
    return M.wigley_preset(name)


@functools.lru_cache(maxsize=None)
def oracle_assembly(kind: str):
    """(mesh, dn_ptr, dn_idx, H, G) for a named test mesh, assembled by the oracle."""
    if kind == "wigley_small":
        m = wigley_small()
    elif kind == "box6":
        m = M.box_mesh(n=6, dirichlet_top=True)
    elif kind == "sphere4":
        m = M.cube_sphere_mesh(n=4)
    elif kind == "tri_block":
        m = M.tri_block_mesh(n=5)
    elif kind == "mixed":
        # a structured box (strip kernel) beside three blocks meeting at a valence-3
        # vertex (no structured grid: column-tile kernel), one system
        m = M.merge_meshes(M.box_mesh(n=5), M.tri_block_mesh(n=4, centre=(3.0, 0, 0)))
    elif kind.startswith("wigley_"):
        m = wigley_preset(kind.split("_", 1)[1])
    else:
        raise KeyError(kind)
    ptr, idx = M.double_nodes_csr(m.support_points)
    H, G = O.assemble(m.support_points, m.cell_dofs, m.cell_orient, ptr, idx)
    return m, ptr, idx, H, G


def explicit_system(H, G, surface, other, cons, tmp_rhs):
    """Dense constrained operator and rhs (ConstrainedOperator semantics)."""
    n = H.shape[0]
    alpha = -H.sum(axis=1)
    A = H * other[None, :] - G * surface[None, :]
    A[np.arange(n), np.arange(n)] += alpha * other
    b = -(H * surface[None, :]) @ tmp_rhs - alpha * surface * tmp_rhs + (G * other[None, :]) @ tmp_rhs
    for l, r in enumerate(cons.rows):
        A[r, :] = 0.0
        A[r, r] = 1.0
        for p in range(cons.ptr[l], cons.ptr[l + 1]):
            A[r, cons.cols[p]] -= cons.weights[p]
        b[r] = cons.inhom[l]
    return A, b


def dense_solution(H, G, surface, other, cons, tmp_rhs):
    A, b = explicit_system(H, G, surface, other, cons, tmp_rhs)
    return np.linalg.solve(A, b)


def row_scaled_err(A, B):
    """max_ij |A_ij - B_ij| / max_j |B_ij| over rows."""
    scale = np.abs(B).max(axis=1)
    scale[scale == 0] = 1.0
    return float((np.abs(A - B).max(axis=1) / scale).max())
