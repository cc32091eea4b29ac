"""Python mirror of WaveBEM's ``BEMProblem<3>`` over the MI355X C-ABI.

Same method names, argument meaning and error behaviour as the reference class
(include/bem_problem.h:87-180): vectors are numpy float64 arrays in the caller's dof
numbering, ``solve``/``solve_system`` overwrite only the unknown entries of ``phi`` and
``dphi_dn`` (source/bem_problem.cc:869-879), and GMRES non-convergence raises
``NoConvergence`` like deal.II's ``SolverControl::NoConvergence``.

Every call goes through ``libwavebem_bem.so``; nothing here computes BEM arithmetic.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .mesh import SurfaceMesh


class WaveBemError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{L.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class NoConvergence(WaveBemError):
    pass


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} entries, got {a.size}")
    return a


@dataclass
class Config:
    """Parameters of ``BEMProblem::parse_parameters`` / ``ComputationalDomain``
    ("Quadrature rules", "Solver"), plus the device-side choices."""
    device: int = 0
    quadrature_order: int = 4
    singular_quadrature_order: int = 5
    solver_tolerance: float = 1e-10
    solver_max_steps: int = 1000
    gmres_max_tmp_vectors: int = 100
    precision: str = "fp64"  # "fp64" | "fp32_ir"
    preconditioner: str = "block_jacobi"  # "block_jacobi" | "none"
    precond_block: int = 64
    ir_max_sweeps: int = 10
    ir_inner_reduction: float = 1e-7
    assembly_tile: int = 16
    ir_inner_format: int = 2  # 0 FP32 hi, 1 row-scaled FP16 copy, 2 compressed FP16 copy
    ir_inner16_reduction: float = 3e-4
    precond_coarse: int = 128  # coarse-space unknowns of the block-Jacobi preconditioner
    assembly_strips: int = 1  # strip sweeps on structured patches (0: column tiles only)
    ir_inner_lr_tol: float = 1e-4  # low-rank tolerance of the compressed inner operator

    def to_c(self) -> L.WbConfig:
        c = L.WbConfig()
        L.lib().wb_config_default(ctypes.byref(c))
        c.device = self.device
        c.quadrature_order = self.quadrature_order
        c.singular_quadrature_order = self.singular_quadrature_order
        c.solver_tolerance = self.solver_tolerance
        c.solver_max_steps = self.solver_max_steps
        c.gmres_max_tmp_vectors = self.gmres_max_tmp_vectors
        c.precision = {"fp64": L.WB_PREC_FP64, "fp32_ir": L.WB_PREC_FP32_IR}[self.precision]
        c.preconditioner = {"none": L.WB_PRECOND_NONE,
                            "block_jacobi": L.WB_PRECOND_BLOCK_JACOBI}[self.preconditioner]
        c.precond_block = self.precond_block
        c.ir_max_sweeps = self.ir_max_sweeps
        c.ir_inner_reduction = self.ir_inner_reduction
        c.assembly_tile = self.assembly_tile
        c.ir_inner_format = self.ir_inner_format
        c.ir_inner16_reduction = self.ir_inner16_reduction
        c.precond_coarse = self.precond_coarse
        c.assembly_strips = self.assembly_strips
        c.ir_inner_lr_tol = self.ir_inner_lr_tol
        return c


class BEMProblem:
    """``BEMProblem<3>`` with the Direct solution method, on one GPU."""

    def __init__(self, mesh: SurfaceMesh | None = None, config: Config | None = None,
                 double_nodes=None):
        self.config = config or Config()
        self._ctx = ctypes.c_void_p()
        cfg = self.config.to_c()
        self._check(L.lib().wb_create(ctypes.byref(self._ctx), ctypes.byref(cfg)), "wb_create")
        self.n = 0
        self.last_stats = None
        if mesh is not None:
            self.reinit(mesh, double_nodes)

    # ---------------------------------------------------------------- plumbing
    def _check(self, status, what):
        if status == L.WB_OK:
            return
        msg = what
        if self._ctx:
            err = L.lib().wb_last_error(self._ctx)
            if err:
                msg = f"{what}: {err.decode()}"
        if status == L.WB_E_NOCONV:
            raise NoConvergence(status, msg)
        raise WaveBemError(status, msg)

    def close(self):
        if self._ctx:
            L.lib().wb_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---------------------------------------------------------------- reference API
    def reinit(self, mesh: SurfaceMesh, double_nodes=None):
        """``BEMProblem::reinit`` plus the ComputationalDomain state it reads."""
        n, nc = mesh.n_dofs, mesh.n_cells
        X = _f64(mesh.support_points).reshape(-1)
        cd = np.ascontiguousarray(mesh.cell_dofs, dtype=np.int32).reshape(-1)
        orient = np.ascontiguousarray(mesh.cell_orient, dtype=np.int8)
        mat = np.ascontiguousarray(mesh.cell_material, dtype=np.int32)
        if double_nodes is not None:
            dp = np.ascontiguousarray(double_nodes[0], dtype=np.int32)
            di = np.ascontiguousarray(double_nodes[1], dtype=np.int32)
            dpp, dip = _i(dp), _i(di)
        else:
            dpp = dip = None
        self._check(L.lib().wb_set_mesh(self._ctx, n, nc, _d(X), _i(cd),
                                        orient.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                                        _i(mat), dpp, dip), "wb_set_mesh")
        self.n = n
        self.mesh = mesh
        self.set_masks(mesh.surface_nodes, mesh.other_nodes)

    def set_masks(self, surface_nodes, other_nodes):
        s = _f64(surface_nodes, self.n)
        o = _f64(other_nodes, self.n)
        self._check(L.lib().wb_set_masks(self._ctx, _d(s), _d(o)), "wb_set_masks")
        self.surface_nodes, self.other_nodes = s.copy(), o.copy()

    def set_support_points(self, support_points):
        """Moves the mesh nodes (same cells, dofs and double nodes); re-assemble afterwards."""
        sp = _f64(support_points, 3 * self.n)
        self._check(L.lib().wb_set_support_points(self._ctx, _d(sp)), "wb_set_support_points")

    def assemble_system(self):
        self._check(L.lib().wb_assemble(self._ctx), "wb_assemble")

    def release_matrices(self):
        """Frees the device G and H (and derived copies) until the next assemble_system."""
        self._check(L.lib().wb_release_matrices(self._ctx), "wb_release_matrices")

    def matrices(self):
        """(G, H) = (dirichlet_matrix, neumann_matrix), caller numbering, float64."""
        G = np.empty((self.n, self.n))
        H = np.empty((self.n, self.n))
        self._check(L.lib().wb_get_matrices(self._ctx, _d(G), _d(H)), "wb_get_matrices")
        return G, H

    def matrix_rows(self, rows):
        """(G[rows], H[rows]) for a list of caller-numbered rows, float64."""
        r = np.ascontiguousarray(rows, dtype=np.int32)
        G = np.empty((r.size, self.n))
        H = np.empty((r.size, self.n))
        self._check(L.lib().wb_get_matrix_rows(self._ctx, int(r.size), _i(r), _d(G), _d(H)),
                    "wb_get_matrix_rows")
        return G, H

    def inner_vmult(self, src):
        """The inner Krylov operator of an FP32-IR context applied to src (FP32 vectors)."""
        s = _f64(src, self.n)
        dst = np.empty(self.n)
        self._check(L.lib().wb_inner_vmult(self._ctx, _d(dst), _d(s)), "wb_inner_vmult")
        return dst

    def inner_operator_info(self):
        """Blocks and bytes of the compressed inner operator (ir_inner_format 2)."""
        info = (ctypes.c_int64 * 6)()
        self._check(L.lib().wb_inner_operator_info(self._ctx, info), "wb_inner_operator_info")
        keys = ("clusters", "dense_blocks", "lowrank_blocks", "zero_blocks", "rank_sum",
                "bytes_per_product")
        return dict(zip(keys, [int(v) for v in info]))

    def coarse_matrix(self):
        """(Ac, agg): the coarse operator P^T A P of the last preconditioner build
        (device dof order aggregates of agg dofs), for inspection."""
        nc, agg = ctypes.c_int(), ctypes.c_int()
        self._check(L.lib().wb_get_coarse_matrix(self._ctx, ctypes.byref(nc), ctypes.byref(agg),
                                                 None), "wb_get_coarse_matrix")
        Ac = np.empty((nc.value, nc.value))
        self._check(L.lib().wb_get_coarse_matrix(self._ctx, ctypes.byref(nc), ctypes.byref(agg),
                                                 _d(Ac)), "wb_get_coarse_matrix")
        return Ac, agg.value

    @property
    def dirichlet_matrix(self):
        return self.matrices()[0]

    @property
    def neumann_matrix(self):
        return self.matrices()[1]

    def compute_alpha(self):
        a = np.empty(self.n)
        self._check(L.lib().wb_compute_alpha(self._ctx, _d(a)), "wb_compute_alpha")
        self.alpha = a
        return a

    def vmult(self, src):
        src = _f64(src, self.n)
        dst = np.empty(self.n)
        self._check(L.lib().wb_vmult(self._ctx, _d(dst), _d(src)), "wb_vmult")
        return dst

    def compute_rhs(self, src):
        src = _f64(src, self.n)
        dst = np.empty(self.n)
        self._check(L.lib().wb_compute_rhs(self._ctx, _d(dst), _d(src)), "wb_compute_rhs")
        return dst

    def compute_normals(self, normalize=True):
        """ComputationalDomain::compute_normals (normalize=True) or
        NumericalTowingTank::compute_normals_at_nodes (normalize=False): (n, 3) L2
        projection of the cell normals at the current support points, on the device."""
        out = np.empty((self.n, 3))
        it = ctypes.c_int(0)
        self._check(L.lib().wb_compute_normals(self._ctx, int(bool(normalize)), _d(out),
                                               ctypes.byref(it)), "wb_compute_normals")
        self.last_projection_iterations = it.value
        return out

    def compute_surface_gradients(self, tmp_rhs):
        """BEMProblem::compute_surface_gradients: (n, 3) L2 projection of the surface
        gradient of tmp_rhs * surface_nodes, on the device."""
        t = _f64(tmp_rhs, self.n)
        out = np.empty((self.n, 3))
        it = ctypes.c_int(0)
        self._check(L.lib().wb_compute_surface_gradients(self._ctx, _d(t), _d(out),
                                                         ctypes.byref(it)),
                    "wb_compute_surface_gradients")
        self.last_projection_iterations = it.value
        return out

    def compute_potential_gradients(self, phi, dphi_dn):
        """FreeSurface::compute_potential_gradients: (n, 3) L2 projection of the complete
        gradient n dphi_dn + surface gradient of phi, averaged over coincident dofs."""
        out = np.empty((self.n, 3))
        it = ctypes.c_int(0)
        self._check(L.lib().wb_compute_potential_gradients(
            self._ctx, _d(_f64(phi, self.n)), _d(_f64(dphi_dn, self.n)), _d(out),
            ctypes.byref(it)), "wb_compute_potential_gradients")
        self.last_projection_iterations = it.value
        return out

    def prepare_bem_vectors(self, free_surface_ids, wall_ids, inflow_ids, wind, bem_bc,
                            dphi_dn, inflow_dphi_dn=None, transom=None, prev_dphi_dn=None):
        """FreeSurface::prepare_bem_vectors: installs the masks of the free-surface cells
        and fills bem_bc / dphi_dn (updated in place) on the other dofs; ``wind`` is the
        (n, 3) wind velocity at the support points."""
        fs = np.ascontiguousarray(free_surface_ids, dtype=np.int32)
        wl = np.ascontiguousarray(wall_ids, dtype=np.int32)
        inf = np.ascontiguousarray(inflow_ids, dtype=np.int32)
        w = _f64(wind, 3 * self.n)
        self._inplace_ok(bem_bc)
        self._inplace_ok(dphi_dn)
        idp = _f64(inflow_dphi_dn, self.n) if inflow_dphi_dn is not None else None
        tr = (np.ascontiguousarray(transom, dtype=np.uint8) if transom is not None else None)
        pv = _f64(prev_dphi_dn, self.n) if prev_dphi_dn is not None else None
        self._check(L.lib().wb_prepare_bem_vectors(
            self._ctx, fs.size, _i(fs), wl.size, _i(wl), inf.size, _i(inf), _d(w),
            _d(idp) if idp is not None else None,
            tr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)) if tr is not None else None,
            _d(pv) if pv is not None else None, _d(bem_bc), _d(dphi_dn)),
            "wb_prepare_bem_vectors")
        s, o = np.empty(self.n), np.empty(self.n)
        self._check(L.lib().wb_get_masks(self._ctx, _d(s), _d(o)), "wb_get_masks")
        self.surface_nodes, self.other_nodes = s, o

    def compute_constraints(self, tmp_rhs, node_normals=None, node_surface_gradients=None):
        t = _f64(tmp_rhs, self.n)
        nn = _f64(node_normals).reshape(-1) if node_normals is not None else None
        gg = _f64(node_surface_gradients).reshape(-1) if node_surface_gradients is not None else None
        nl = ctypes.c_int(0)
        self._check(L.lib().wb_compute_constraints(
            self._ctx, _d(t), _d(nn) if nn is not None else None,
            _d(gg) if gg is not None else None, ctypes.byref(nl)), "wb_compute_constraints")
        return nl.value

    def set_constraints(self, rows, ptr, cols, weights, inhom):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        ptr = np.ascontiguousarray(ptr, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        w = _f64(weights)
        inh = _f64(inhom)
        self._check(L.lib().wb_set_constraints(self._ctx, rows.size, _i(rows), _i(ptr), _i(cols),
                                               _d(w), _d(inh)), "wb_set_constraints")

    def constrained_vmult(self, src):
        src = _f64(src, self.n)
        dst = np.empty(self.n)
        self._check(L.lib().wb_constrained_vmult(self._ctx, _d(dst), _d(src)),
                    "wb_constrained_vmult")
        return dst

    def solve_system(self, phi, dphi_dn, tmp_rhs):
        """Overwrites the unknown entries of phi / dphi_dn in place."""
        t = _f64(tmp_rhs, self.n)
        self._inplace_ok(phi)
        self._inplace_ok(dphi_dn)
        st = L.WbSolveStats()
        status = L.lib().wb_solve_system(self._ctx, _d(t), _d(phi), _d(dphi_dn), ctypes.byref(st))
        self.last_stats = st.as_dict()
        self._check(status, "wb_solve_system")
        return self.last_stats

    def solve(self, phi, dphi_dn, tmp_rhs):
        """``BEMProblem::solve`` (Direct branch): assemble_system + solve_system in one call
        (wb_solve); phi / dphi_dn are updated in place."""
        t = _f64(tmp_rhs, self.n)
        self._inplace_ok(phi)
        self._inplace_ok(dphi_dn)
        st = L.WbSolveStats()
        status = L.lib().wb_solve(self._ctx, _d(t), _d(phi), _d(dphi_dn), ctypes.byref(st))
        self.last_stats = st.as_dict()
        self._check(status, "wb_solve")
        return self.last_stats

    def residual(self, phi, dphi_dn):
        p = _f64(phi, self.n)
        d = _f64(dphi_dn, self.n)
        res = np.empty(self.n)
        self._check(L.lib().wb_residual(self._ctx, _d(res), _d(p), _d(d)), "wb_residual")
        return res

    def solve_system_multi(self, phi, dphi_dn, tmp_rhs):
        """k right-hand sides as (k, n) arrays; phi/dphi_dn updated in place."""
        t = _f64(tmp_rhs)
        k = t.size // self.n
        for a in (phi, dphi_dn):
            self._inplace_ok(a)
            if a.size != k * self.n:
                raise ValueError("phi/dphi_dn must have shape (k, n)")
        st = L.WbSolveStats()
        status = L.lib().wb_solve_system_multi(self._ctx, k, _d(t), _d(phi), _d(dphi_dn), None,
                                               ctypes.byref(st))
        self.last_stats = st.as_dict()
        self._check(status, "wb_solve_system_multi")
        return self.last_stats

    def solve_adjoint(self, rhs):
        """lambda with A_c^T lambda = rhs (transposed constrained operator), float64 (n,)."""
        r = _f64(rhs, self.n)
        lam = np.zeros(self.n)
        st = L.WbSolveStats()
        status = L.lib().wb_solve_adjoint(self._ctx, _d(r), _d(lam), ctypes.byref(st))
        self.last_stats = st.as_dict()
        self._check(status, "wb_solve_adjoint")
        return lam

    def residual_vjp(self, w, phi, dphi_dn, free_rows_only=True):
        """(n, 3): sum_i w_i dR_i/dx_d, R_i = (alpha_i phi_i + sum_j H_ij phi_j)
        - sum_j G_ij dphi_dn_j, at the current assembly."""
        g = np.zeros((self.n, 3))
        self._check(L.lib().wb_residual_vjp(self._ctx, _d(_f64(w, self.n)),
                                            _d(_f64(phi, self.n)), _d(_f64(dphi_dn, self.n)),
                                            int(free_rows_only), _d(g)), "wb_residual_vjp")
        return g

    def residual_jvp(self, dx, phi, dphi_dn, free_rows_only=True):
        """(n,): sum_d dR_i/dx_d . dx_d for a support-point direction dx (n, 3)."""
        out = np.zeros(self.n)
        self._check(L.lib().wb_residual_jvp(self._ctx, _d(_f64(dx, 3 * self.n)),
                                            _d(_f64(phi, self.n)), _d(_f64(dphi_dn, self.n)),
                                            int(free_rows_only), _d(out)), "wb_residual_jvp")
        return out

    def solve_rhs(self, rhs):
        """sol with A_c sol = rhs (the solve_system operator, a given right-hand side)."""
        sol = np.zeros(self.n)
        st = L.WbSolveStats()
        status = L.lib().wb_solve_rhs(self._ctx, _d(_f64(rhs, self.n)), _d(sol),
                                      ctypes.byref(st))
        self.last_stats = st.as_dict()
        self._check(status, "wb_solve_rhs")
        return sol

    def solve_system_dev(self, phi_ptr, dphi_dn_ptr, tmp_rhs_ptr):
        """Device-pointer form (caller dof order, n float64 each)."""
        st = L.WbSolveStats()
        status = L.lib().wb_solve_system_dev(self._ctx, ctypes.c_void_p(tmp_rhs_ptr),
                                             ctypes.c_void_p(phi_ptr),
                                             ctypes.c_void_p(dphi_dn_ptr), ctypes.byref(st))
        self.last_stats = st.as_dict()
        self._check(status, "wb_solve_system_dev")
        return self.last_stats

    def set_profiling(self, enable=True):
        self._check(L.lib().wb_set_profiling(self._ctx, int(enable)), "wb_set_profiling")

    def kernel_stats(self):
        ks = L.WbKernelStats()
        self._check(L.lib().wb_get_kernel_stats(self._ctx, ctypes.byref(ks)),
                    "wb_get_kernel_stats")
        return ks.as_dict()

    def device_bytes(self):
        return int(L.lib().wb_device_bytes(self._ctx))

    @staticmethod
    def _inplace_ok(a):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
            raise TypeError("phi / dphi_dn must be C-contiguous float64 numpy arrays")


# ---------------------------------------------------------------- host utilities
def host_quadrature(q_order=4, s_order=5):
    nq, ns = q_order * q_order, 2 * s_order * s_order
    rp, rw = np.zeros((nq, 2)), np.zeros(nq)
    sp, sw = np.zeros((4, ns, 2)), np.zeros((4, ns))
    st = L.lib().wb_host_quadrature(q_order, s_order, _d(rp), _d(rw), _d(sp), _d(sw))
    if st != L.WB_OK:
        raise WaveBemError(st, "wb_host_quadrature")
    return rp, rw, sp, sw


def host_strip_plan(mesh: SurfaceMesh, strips: int = 1):
    """The strip schedule of the regular-cell assembly (wb_host_strip_plan)."""
    n, nc = mesh.n_dofs, mesh.n_cells
    X = _f64(mesh.support_points).reshape(-1)
    cd = np.ascontiguousarray(mesh.cell_dofs, dtype=np.int32).reshape(-1)
    mat = np.ascontiguousarray(mesh.cell_material, dtype=np.int32)
    sizes = np.zeros(5, dtype=np.int32)
    st = L.lib().wb_host_strip_plan(n, nc, _d(X), _i(cd), _i(mat), strips, _i(sizes), None, None,
                                    None, None, None)
    if st != L.WB_OK:
        raise WaveBemError(st, "wb_host_strip_plan")
    nt, ng = int(sizes[0]), int(sizes[1])
    tasks = np.zeros((max(nt, 1), 2), dtype=np.int32)
    gaps = np.zeros((max(ng, 1), 4), dtype=np.int32)
    gcells = np.zeros((max(ng, 1), 9), dtype=np.int32)
    cis = np.zeros((nc, 4), dtype=np.int32)
    perm = np.zeros(n, dtype=np.int32)
    st = L.lib().wb_host_strip_plan(n, nc, _d(X), _i(cd), _i(mat), strips, _i(sizes), _i(tasks),
                                    _i(gaps), _i(gcells), _i(cis), _i(perm))
    if st != L.WB_OK:
        raise WaveBemError(st, "wb_host_strip_plan")
    return dict(tasks=tasks[:nt], gaps=gaps[:ng], gap_cells=gcells[:ng], cell_int_s=cis,
                perm=perm, cells=int(sizes[2]), strip_dofs=int(sizes[3]), n_tiles=int(sizes[4]))


def host_cluster_plan(mesh: SurfaceMesh):
    """The cluster plan of the compressed inner operator (wb_host_cluster_plan)."""
    n, nc = mesh.n_dofs, mesh.n_cells
    X = _f64(mesh.support_points).reshape(-1)
    cd = np.ascontiguousarray(mesh.cell_dofs, dtype=np.int32).reshape(-1)
    mat = np.ascontiguousarray(mesh.cell_material, dtype=np.int32)
    ncl = ctypes.c_int32()
    st = L.lib().wb_host_cluster_plan(n, nc, _d(X), _i(cd), _i(mat), ctypes.byref(ncl), None,
                                      None, None, None)
    if st != L.WB_OK:
        raise WaveBemError(st, "wb_host_cluster_plan")
    k = ncl.value
    size = np.zeros(k, dtype=np.int32)
    idx = np.zeros((k, 256), dtype=np.int32)
    near = np.zeros((k, k), dtype=np.uint8)
    perm = np.zeros(n, dtype=np.int32)
    st = L.lib().wb_host_cluster_plan(n, nc, _d(X), _i(cd), _i(mat), ctypes.byref(ncl), _i(size),
                                      _i(idx), near.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                      _i(perm))
    if st != L.WB_OK:
        raise WaveBemError(st, "wb_host_cluster_plan")
    return dict(size=size, idx=idx, near=near.astype(bool), perm=perm)


def host_hm_work_items(rank, cl_size):
    """The compressed product's work lists for block ranks (ncl x ncl uint8: 0 zero, 255
    dense, else the rank) and cluster sizes (wb_host_hm_work_items): dict with chunks,
    item_ptr, items, dcols, shift, lowrank."""
    r = np.ascontiguousarray(rank, dtype=np.uint8)
    ncl = r.shape[0]
    sz = np.ascontiguousarray(cl_size, dtype=np.int32)
    sizes = np.zeros(5, dtype=np.int32)
    rp = r.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    st = L.lib().wb_host_hm_work_items(ncl, rp, _i(sz), _i(sizes), None, None)
    if st != L.WB_OK:
        raise WaveBemError(st, "wb_host_hm_work_items")
    ptr = np.zeros(ncl * int(sizes[0]) + 1, dtype=np.int32)
    items = np.zeros(max(int(sizes[1]), 1), dtype=np.int32)
    st = L.lib().wb_host_hm_work_items(ncl, rp, _i(sz), _i(sizes), _i(ptr), _i(items))
    if st != L.WB_OK:
        raise WaveBemError(st, "wb_host_hm_work_items")
    return dict(chunks=int(sizes[0]), item_ptr=ptr, items=items[:int(sizes[1])],
                dcols=int(sizes[2]), shift=int(sizes[3]), lowrank=int(sizes[4]))


def host_mesh_plan(mesh: SurfaceMesh):
    n, nc = mesh.n_dofs, mesh.n_cells
    X = _f64(mesh.support_points).reshape(-1)
    cd = np.ascontiguousarray(mesh.cell_dofs, dtype=np.int32).reshape(-1)
    mat = np.ascontiguousarray(mesh.cell_material, dtype=np.int32)
    perm = np.zeros(n, dtype=np.int32)
    dp = np.zeros(n + 1, dtype=np.int32)
    cap = 8 * n
    di = np.zeros(cap, dtype=np.int32)
    nt, ntc, nsp = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    st = L.lib().wb_host_mesh_plan(n, nc, _d(X), _i(cd), _i(mat), _i(perm), _i(dp), _i(di), cap,
                                   ctypes.byref(nt), ctypes.byref(ntc), ctypes.byref(nsp))
    if st != L.WB_OK:
        raise WaveBemError(st, "wb_host_mesh_plan")
    return dict(perm=perm, dn_ptr=dp, dn_idx=di[:dp[-1]].copy(), n_tiles=nt.value,
                n_tile_cells=ntc.value, n_singular_pairs=nsp.value)


def host_compute_constraints(dn_ptr, dn_idx, surface, tmp_rhs, normals=None, grads=None):
    dp = np.ascontiguousarray(dn_ptr, dtype=np.int32)
    di = np.ascontiguousarray(dn_idx, dtype=np.int32)
    n = dp.size - 1
    cap_l, cap_nnz = n, 8 * n + 16
    rows = np.zeros(cap_l, dtype=np.int32)
    ptr = np.zeros(cap_l + 1, dtype=np.int32)
    cols = np.zeros(cap_nnz, dtype=np.int32)
    w = np.zeros(cap_nnz)
    inh = np.zeros(cap_l)
    nn = _f64(normals).reshape(-1) if normals is not None else None
    gg = _f64(grads).reshape(-1) if grads is not None else None
    nl = L.lib().wb_host_compute_constraints(
        n, _i(dp), _i(di), _d(_f64(surface, n)), _d(nn) if nn is not None else None,
        _d(gg) if gg is not None else None, _d(_f64(tmp_rhs, n)), cap_l, cap_nnz, _i(rows),
        _i(ptr), _i(cols), _d(w), _d(inh))
    if nl < 0:
        raise WaveBemError(-nl, "wb_host_compute_constraints")
    nnz = int(ptr[nl])
    return dict(rows=rows[:nl].copy(), ptr=ptr[:nl + 1].copy(), cols=cols[:nnz].copy(),
                weights=w[:nnz].copy(), inhom=inh[:nl].copy())
