"""Python face of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import
this package, as the checker.  The product library never touches it.

Every function wraps the C restatement in ``bem_oracle.c``; see that file for the
WaveBEM file:line each routine follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_bp = ctypes.POINTER(ctypes.c_byte)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def qgauss_1d(n):
    x = np.zeros(n)
    w = np.zeros(n)
    assert lib().orc_qgauss_1d(n, _d(x), _d(w)) == 0
    return x, w


def qgauss_2d(n):
    p = np.zeros((n * n, 2))
    w = np.zeros(n * n)
    assert lib().orc_qgauss_2d(n, _d(p), _d(w)) == 0
    return p, w


def qgauss_one_over_r(n, vertex, factor_out=True):
    p = np.zeros((2 * n * n, 2))
    w = np.zeros(2 * n * n)
    assert lib().orc_qgauss_one_over_r(n, vertex, int(factor_out), _d(p), _d(w)) == 0
    return p, w


def fe_values(vertices, orient, pts, w):
    v = _f64(vertices).reshape(12)
    pts = _f64(pts)
    w = _f64(w)
    nq = w.size
    qp = np.zeros((nq, 3))
    nr = np.zeros((nq, 3))
    jxw = np.zeros(nq)
    sh = np.zeros((nq, 4))
    lib().orc_fe_values(_d(v), int(orient), nq, _d(pts), _d(w), _d(qp), _d(nr), _d(jxw), _d(sh))
    return qp, nr, jxw, sh


def assemble(support, cell_dofs, cell_orient, dn_ptr, dn_idx, q_order=4, sing_order=5,
             n_threads=0):
    """Returns (H, G) = (neumann_matrix, dirichlet_matrix)."""
    support = _f64(support)
    cd = _i32(cell_dofs)
    orient = np.ascontiguousarray(cell_orient, dtype=np.int8)
    dp = _i32(dn_ptr)
    di = _i32(dn_idx)
    n = support.shape[0]
    nc = cd.shape[0]
    H = np.zeros((n, n))
    G = np.zeros((n, n))
    rc = lib().orc_assemble(n, nc, _i(cd), orient.ctypes.data_as(_bp), _d(support), _i(dp),
                            _i(di), q_order, sing_order, _d(H), _d(G), int(n_threads))
    if rc != 0:
        raise RuntimeError(f"orc_assemble failed: {rc}")
    return H, G


def assemble_rows(support, cell_dofs, cell_orient, dn_ptr, dn_idx, row0, row1, q_order=4,
                  sing_order=5, n_threads=0):
    """Rows [row0, row1) of (H, G)."""
    support = _f64(support)
    cd = _i32(cell_dofs)
    orient = np.ascontiguousarray(cell_orient, dtype=np.int8)
    dp = _i32(dn_ptr)
    di = _i32(dn_idx)
    n = support.shape[0]
    nc = cd.shape[0]
    H = np.zeros((row1 - row0, n))
    G = np.zeros((row1 - row0, n))
    rc = lib().orc_assemble_rows(n, nc, _i(cd), orient.ctypes.data_as(_bp), _d(support), _i(dp),
                                 _i(di), q_order, sing_order, int(row0), int(row1), _d(H), _d(G),
                                 int(n_threads))
    if rc != 0:
        raise RuntimeError(f"orc_assemble_rows failed: {rc}")
    return H, G


def residual_vjp(support, cell_dofs, cell_orient, dn_ptr, dn_idx, w, phi, psi, q_order=4,
                 sing_order=5):
    """(n, 3): sum_i w_i dR_i/dx_d, R_i = sum_j H_ij (phi_j - phi_i) - sum_j G_ij psi_j."""
    support = _f64(support)
    cd = _i32(cell_dofs)
    orient = np.ascontiguousarray(cell_orient, dtype=np.int8)
    n = support.shape[0]
    out = np.zeros((n, 3))
    rc = lib().orc_residual_vjp(n, cd.shape[0], _i(cd), orient.ctypes.data_as(_bp),
                                _d(support), _i(_i32(dn_ptr)), _i(_i32(dn_idx)), q_order,
                                sing_order, _d(_f64(w)), _d(_f64(phi)), _d(_f64(psi)), _d(out))
    if rc != 0:
        raise RuntimeError(f"orc_residual_vjp failed: {rc}")
    return out


def node_normals(support, cell_dofs, cell_orient, q_order=4, normalize=True):
    """(n, 3) nodal normals: L2 projection of the cell normals, normalised per node unless
    ``normalize`` is False."""
    support = _f64(support)
    cd = _i32(cell_dofs)
    orient = np.ascontiguousarray(cell_orient, dtype=np.int8)
    out = np.zeros((support.shape[0], 3))
    rc = lib().orc_node_normals(support.shape[0], cd.shape[0], _i(cd),
                                orient.ctypes.data_as(_bp), _d(support), q_order,
                                int(bool(normalize)), _d(out))
    if rc != 0:
        raise RuntimeError(f"orc_node_normals failed: {rc}")
    return out


def surface_gradients(support, cell_dofs, cell_orient, phi, q_order=4):
    """(n, 3) L2 projection of the surface gradient of the Q1 interpolant of phi."""
    support = _f64(support)
    cd = _i32(cell_dofs)
    orient = np.ascontiguousarray(cell_orient, dtype=np.int8)
    out = np.zeros((support.shape[0], 3))
    rc = lib().orc_surface_gradients(support.shape[0], cd.shape[0], _i(cd),
                                     orient.ctypes.data_as(_bp), _d(support), q_order,
                                     _d(_f64(phi)), _d(out))
    if rc != 0:
        raise RuntimeError(f"orc_surface_gradients failed: {rc}")
    return out


def potential_gradients(support, cell_dofs, cell_orient, phi, psi, dn_ptr, dn_idx, q_order=4,
                        average=True):
    """FreeSurface::compute_potential_gradients (source/free_surface.cc:9271-9400), restated
    in numpy: on every cell and point of the QGauss rule, the complete gradient
    g = n psi + surface gradient of the Q1 interpolant of phi (:9325-9344); the L2 system
    sum_j M_ij x_j = sum_c sum_q N_i g JxW, one scalar mass matrix for the three components
    (the vector mass matrix of :9346-9358 couples equal components only), solved directly
    (the reference's SparseDirectUMFPACK, :9374-9376); then every set of coincident dofs
    takes its mean (:9380-9396).  Returns (n, 3)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    X = _f64(support)
    cd = np.asarray(cell_dofs, dtype=np.int64)
    o = np.asarray(cell_orient, dtype=np.float64)
    n = X.shape[0]
    pts, w = qgauss_2d(q_order)
    xi, eta = pts[:, 0], pts[:, 1]
    sh = np.stack([(1 - xi) * (1 - eta), xi * (1 - eta), (1 - xi) * eta, xi * eta], 1)
    dxi = np.stack([-(1 - eta), (1 - eta), -eta, eta], 1)
    deta = np.stack([-(1 - xi), -xi, (1 - xi), xi], 1)
    V = X[cd]                                            # (c, 4, 3)
    a = np.einsum("qv,cvd->cqd", dxi, V)
    b = np.einsum("qv,cvd->cqd", deta, V)
    cr = np.cross(a, b)
    ln = np.linalg.norm(cr, axis=2)
    nrm = o[:, None, None] * cr / ln[:, :, None]
    g00, g01, g11 = (a * a).sum(2), (a * b).sum(2), (b * b).sum(2)
    det = g00 * g11 - g01 * g01
    jxw = np.sqrt(det) * w[None, :]
    pc = np.asarray(phi, dtype=np.float64)[cd]
    ux, ue = pc @ dxi.T, pc @ deta.T
    ca = (g11 * ux - g01 * ue) / det
    cb = (g00 * ue - g01 * ux) / det
    psq = np.asarray(psi, dtype=np.float64)[cd] @ sh.T
    grad = ca[:, :, None] * a + cb[:, :, None] * b + nrm * psq[:, :, None]
    Me = np.einsum("qi,qj,cq->cij", sh, sh, jxw)
    re = np.einsum("qi,cqd,cq->cid", sh, grad, jxw)
    rows = np.repeat(cd, 4, axis=1).ravel()
    cols = np.tile(cd, (1, 4)).ravel()
    M = sp.csc_matrix((Me.ravel(), (rows, cols)), shape=(n, n))
    rhs = np.zeros((n, 3))
    np.add.at(rhs, cd.ravel(), re.reshape(-1, 3))
    lu = spla.splu(M)
    out = np.stack([lu.solve(rhs[:, e]) for e in range(3)], 1)
    if average:
        ptr, idx = np.asarray(dn_ptr), np.asarray(dn_idx)
        avg = out.copy()
        for i in range(n):
            if ptr[i + 1] - ptr[i] > 1:
                avg[i] = out[idx[ptr[i]:ptr[i + 1]]].mean(axis=0)
        out = avg
    return out


def prepare_bem_vectors(support, cell_dofs, cell_orient, cell_material, dn_ptr, dn_idx,
                        fs_ids, wall_ids, inflow_ids, wind, bem_bc, dphi_dn,
                        inflow_dphi_dn=None, transom=None, prev_dphi_dn=None, q_order=4):
    """FreeSurface::prepare_bem_vectors (source/free_surface.cc:8744-8883).  Returns
    (surface_nodes, other_nodes, bem_bc, dphi_dn); the inputs are not modified.
      * surface_nodes = 1 on the dofs of free-surface cells (:8753-8781);
      * on the dofs of the other cells, in cell order (:8787-8847): wall cells
        -node_normals . wind with the unnormalised projected normals
        (NumericalTowingTank::compute_normals_at_nodes, numerical_towing_tank.cc:3358-3433),
        inflow cells inflow_dphi_dn, others 0, written to both vectors;
      * transom dofs (:8856-8880): Neumann, mean of prev_dphi_dn over the other dofs of
        their double-node set."""
    n = support.shape[0]
    surf = np.zeros(n)
    oth = np.ones(n)
    bc = np.array(bem_bc, dtype=np.float64)
    dn = np.array(dphi_dn, dtype=np.float64)
    for k in range(cell_dofs.shape[0]):
        if cell_material[k] in fs_ids:
            surf[cell_dofs[k]] = 1.0
            oth[cell_dofs[k]] = 0.0
    nn = node_normals(support, cell_dofs, cell_orient, q_order, normalize=False)
    for k in range(cell_dofs.shape[0]):
        m = cell_material[k]
        if m in fs_ids:
            continue
        for j in cell_dofs[k]:
            if m in wall_ids:
                v = -(nn[j, 0] * wind[j, 0] + nn[j, 1] * wind[j, 1] + nn[j, 2] * wind[j, 2])
            elif m in inflow_ids:
                v = 0.0 if inflow_dphi_dn is None else inflow_dphi_dn[j]
            else:
                v = 0.0
            bc[j] = v
            dn[j] = v
    if transom is not None:
        for i in range(n):
            if not transom[i]:
                continue
            surf[i] = 0.0
            oth[i] = 1.0
            dup = [int(d) for d in dn_idx[dn_ptr[i]:dn_ptr[i + 1]] if d != i]
            v = 0.0
            for d in dup:
                v += prev_dphi_dn[d] / len(dup)
            bc[i] = v
            dn[i] = v
    return surf, oth, bc, dn


class FMA:
    """BEMFMA restated (bem_oracle.h orc_fma_*): octree + near field + multipole far field.
    ``products(phi, dphi_dn)`` returns (prodN ~ H phi, prodD ~ G dphi_dn)."""

    def __init__(self, support, cell_dofs, cell_orient, dn_ptr, dn_idx, trunc_order=4,
                 num_levels=10, q_order=4, sing_order=5):
        self._keep = [np.ascontiguousarray(support, dtype=np.float64),
                      np.ascontiguousarray(cell_dofs, dtype=np.int32),
                      np.ascontiguousarray(cell_orient, dtype=np.int8),
                      np.ascontiguousarray(dn_ptr, dtype=np.int32),
                      np.ascontiguousarray(dn_idx, dtype=np.int32)]
        X, cd, orient, dp, di = self._keep
        self.n = X.shape[0]
        L = lib()
        L.orc_fma_create.restype = ctypes.c_void_p
        L.orc_fma_destroy.argtypes = [ctypes.c_void_p]
        L.orc_fma_products.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_fma_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
        self._h = L.orc_fma_create(self.n, cd.shape[0], _i(cd), orient.ctypes.data_as(_bp),
                                   _d(X), _i(dp), _i(di), q_order, sing_order, trunc_order,
                                   num_levels)
        if not self._h:
            raise RuntimeError("orc_fma_create failed")

    def info(self):
        a = (ctypes.c_long * 5)()
        lib().orc_fma_info(self._h, a)
        return dict(blocks=a[0], childless=a[1], near_entries=a[2], elem_expansions=a[3],
                    levels=a[4])

    def coverage(self, n_cells, nq=16):
        """(n, n_cells * nq) uint8: how often each (node, quadrature point) pair is used."""
        c = np.zeros((self.n, n_cells * nq), dtype=np.uint8)
        lib().orc_fma_coverage(ctypes.c_void_p(self._h), c.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
        return c

    def products(self, phi, dphi_dn):
        pn = np.zeros(self.n)
        pd = np.zeros(self.n)
        lib().orc_fma_products(self._h, _d(_f64(phi)), _d(_f64(dphi_dn)), _d(pn), _d(pd))
        return pn, pd

    def close(self):
        if self._h:
            lib().orc_fma_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fma_expansion_check(p, kind, src, centres, x, normal=(0.0, 0.0, 1.0)):
    c = np.zeros(9)
    cc = np.asarray(centres, dtype=np.float64).ravel()
    c[:cc.size] = cc
    out = ctypes.c_double(0.0)
    rc = lib().orc_fma_expansion_check(p, kind, _d(_f64(src)), _d(_f64(normal)), _d(c),
                                       _d(_f64(x)), ctypes.byref(out))
    if rc != 0:
        raise ValueError("bad expansion check arguments")
    return out.value


def fma_legendre(n, m, x):
    lib().orc_fma_legendre.restype = ctypes.c_double
    ds = ctypes.c_double(0.0)
    v = lib().orc_fma_legendre(n, m, ctypes.c_double(x), ctypes.byref(ds))
    return v, ds.value


def gemv_time(n, reps=3, n_threads=0):
    """Seconds per dense FP64 row-major mat-vec of size n (FullMatrix::vmult loop)."""
    import time
    M = np.random.default_rng(0).standard_normal((n, n))
    x = np.ones(n)
    y = np.zeros(n)
    lib().orc_gemv_timed.restype = ctypes.c_double
    return lib().orc_gemv_timed(n, _d(M), _d(x), _d(y), int(reps), int(n_threads))


def compute_alpha(H):
    H = _f64(H)
    n = H.shape[0]
    a = np.zeros(n)
    lib().orc_compute_alpha(n, _d(H), _d(a))
    return a


def vmult(G, H, alpha, surface, other, src):
    n = G.shape[0]
    dst = np.zeros(n)
    lib().orc_vmult(n, _d(_f64(G)), _d(_f64(H)), _d(_f64(alpha)), _d(_f64(surface)),
                    _d(_f64(other)), _d(_f64(src)), _d(dst))
    return dst


def compute_rhs(G, H, alpha, surface, other, src):
    n = G.shape[0]
    dst = np.zeros(n)
    lib().orc_compute_rhs(n, _d(_f64(G)), _d(_f64(H)), _d(_f64(alpha)), _d(_f64(surface)),
                          _d(_f64(other)), _d(_f64(src)), _d(dst))
    return dst


@dataclass
class Constraints:
    rows: np.ndarray
    ptr: np.ndarray
    cols: np.ndarray
    weights: np.ndarray
    inhom: np.ndarray

    @property
    def n_lines(self):
        return int(self.rows.size)


def compute_constraints(dn_ptr, dn_idx, surface, tmp_rhs, normals=None, grads=None):
    dp = _i32(dn_ptr)
    di = _i32(dn_idx)
    n = dp.size - 1
    cap_l = n
    cap_nnz = 8 * n + 16
    rows = np.zeros(cap_l, dtype=np.int32)
    ptr = np.zeros(cap_l + 1, dtype=np.int32)
    cols = np.zeros(cap_nnz, dtype=np.int32)
    w = np.zeros(cap_nnz)
    inh = np.zeros(cap_l)
    nrm = _f64(normals) if normals is not None else None
    grd = _f64(grads) if grads is not None else None
    nl = lib().orc_compute_constraints(
        n, _i(dp), _i(di), _d(_f64(surface)), _d(nrm) if nrm is not None else None,
        _d(grd) if grd is not None else None, _d(_f64(tmp_rhs)), cap_l, cap_nnz, _i(rows),
        _i(ptr), _i(cols), _d(w), _d(inh))
    if nl < 0:
        raise RuntimeError(f"orc_compute_constraints failed: {nl}")
    nnz = int(ptr[nl])
    return Constraints(rows[:nl].copy(), ptr[:nl + 1].copy(), cols[:nnz].copy(), w[:nnz].copy(),
                       inh[:nl].copy())


def constrained_vmult(G, H, alpha, surface, other, cons: Constraints, src):
    n = G.shape[0]
    dst = np.zeros(n)
    lib().orc_constrained_vmult(n, _d(_f64(G)), _d(_f64(H)), _d(_f64(alpha)), _d(_f64(surface)),
                                _d(_f64(other)), cons.n_lines, _i(cons.rows), _i(cons.ptr),
                                _i(cons.cols), _d(cons.weights), _d(_f64(src)), _d(dst))
    return dst


@dataclass
class SolveResult:
    phi: np.ndarray
    dphi_dn: np.ndarray
    alpha: np.ndarray
    sol: np.ndarray
    iterations: int
    residual: float
    status: int


def solve_system(G, H, surface, other, cons: Constraints, tmp_rhs, phi=None, dphi_dn=None,
                 tol=1e-10, max_steps=1000, restart_vectors=100, precond=1, band=100):
    n = G.shape[0]
    phi = np.zeros(n) if phi is None else _f64(phi).copy()
    dphi = np.zeros(n) if dphi_dn is None else _f64(dphi_dn).copy()
    alpha = np.zeros(n)
    sol = np.zeros(n)
    it = ctypes.c_int(0)
    res = ctypes.c_double(0)
    rc = lib().orc_solve_system(
        n, _d(_f64(G)), _d(_f64(H)), _d(_f64(surface)), _d(_f64(other)), cons.n_lines,
        _i(cons.rows), _i(cons.ptr), _i(cons.cols), _d(cons.weights), _d(cons.inhom),
        _d(_f64(tmp_rhs)), ctypes.c_double(tol), int(max_steps), int(restart_vectors),
        int(precond), int(band), _d(phi), _d(dphi), _d(alpha), _d(sol), ctypes.byref(it),
        ctypes.byref(res))
    if rc < 0:
        raise RuntimeError(f"orc_solve_system failed: {rc}")
    return SolveResult(phi, dphi, alpha, sol, it.value, res.value, rc)


def residual(G, H, surface, other, cons: Constraints, phi, dphi_dn):
    n = G.shape[0]
    res = np.zeros(n)
    lib().orc_residual(n, _d(_f64(G)), _d(_f64(H)), _d(_f64(surface)), _d(_f64(other)),
                       cons.n_lines, _i(cons.rows), _i(cons.ptr), _i(cons.cols),
                       _d(cons.weights), _d(cons.inhom), _d(_f64(phi)), _d(_f64(dphi_dn)),
                       _d(res))
    return res


def _setup_argtypes():
    L = lib()
    L.orc_solve_system.argtypes = None  # varargs-free; ctypes converts per call


if __name__ == "__main__":
    build(force=True)
