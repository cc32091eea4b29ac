/*
 * wavebem_bem.h -- C-ABI of the MI355X (gfx950) collocation-BEM library.
 *
 * The entry points replace the arithmetic bodies of WaveBEM's BEMProblem<3>
 * (include/bem_problem.h:87-180, source/bem_problem.cc) behind its unchanged class
 * interface.  Each declaration names the reference member it stands in for.  Plain
 * pointers and sizes only; every function returns a wb_status and never throws.
 *
 * Index conventions: dof indices and vectors are in the caller's (deal.II) dof
 * numbering.  Matrices returned by wb_get_matrices are row-major n x n like deal.II's
 * FullMatrix.  Host pointers unless a function name ends in _dev.
 *
 * Threading: one wb_ctx per host thread; each context owns one HIP stream.  There is no
 * global mutable state (the reference keeps static locals in compute_alpha, vmult,
 * compute_rhs and assemble_preconditioner; here that state lives in the context).
 */
#ifndef WAVEBEM_BEM_H
#define WAVEBEM_BEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    WB_OK = 0,
    WB_E_ARG = 1,       /* invalid argument / inconsistent sizes */
    WB_E_OOM = 2,       /* device or host allocation failed */
    WB_E_NOCONV = 3,    /* SolverControl::NoConvergence (max steps reached / NaN) */
    WB_E_HIP = 4,       /* HIP runtime error */
    WB_E_STATE = 5,     /* call out of order (e.g. solve before set_mesh) */
    WB_E_UNSUPPORTED = 6
} wb_status;

typedef enum {
    WB_PREC_FP64 = 0,    /* G, H stored and applied in FP64 */
    WB_PREC_FP32_IR = 1  /* FP64 quadrature; G, H stored as FP32 + bf16 correction;
                            FP32 GMRES inside FP64 iterative refinement */
} wb_precision;

typedef enum {
    WB_PRECOND_NONE = 0,
    WB_PRECOND_BLOCK_JACOBI = 1  /* stands in for the banded UMFPACK LU of
                                    source/bem_problem.cc:1107-1149 */
} wb_preconditioner;

/* Parameters the reference reads from its .prm file:
 * "Quadrature rules" (source/computational_domain.cc:84-91), "Solver" SolverControl
 * and GMRES AdditionalData(100) (source/bem_problem.cc:83-85,826-827). */
typedef struct {
    int device;                    /* HIP device ordinal */
    int quadrature_order;          /* QGauss<2>(q): 4 */
    int singular_quadrature_order; /* QGaussOneOverR<2>(s): 5 */
    double solver_tolerance;       /* absolute, preconditioned residual: 1e-10 */
    int solver_max_steps;          /* total GMRES iterations: 1000 */
    int gmres_max_tmp_vectors;     /* AdditionalData(100): Krylov dim = value - 2 */
    int precision;                 /* wb_precision */
    int preconditioner;            /* wb_preconditioner */
    int precond_block;             /* block size for block Jacobi (<= 128) */
    int ir_max_sweeps;             /* FP32_IR: outer refinement sweeps */
    double ir_inner_reduction;     /* FP32_IR: an inner FP32 solve stops at
                                      max(this * rho, tol / 2), rho = outer residual */
    int assembly_tile;             /* max columns per assembly column tile (1..16) */
    int ir_inner_format;           /* FP32_IR: the inner Krylov solves run on 0 = the FP32
                                      (hi) operator, 1 = a row-scaled FP16 copy of it (2
                                      bytes per entry), 2 (default) = a compressed copy:
                                      blocks of
                                      clusters of <= 256 dofs, FP16, dense where the
                                      clusters touch, else low rank (Q B, rank <= 12) where
                                      that meets ir_inner_lr_tol.  The refinement residual
                                      keeps the FP32 hi + bf16 lo operator whatever the
                                      format; with 1 or 2 a sweep that fails to halve the
                                      residual switches the remaining sweeps to the FP32
                                      operator */
    double ir_inner16_reduction;   /* FP32_IR with ir_inner_format = 1 or 2: the inner solve
                                      stops at max(this * rho, tol / 2) */
    int precond_coarse;            /* block Jacobi only: at most this many coarse unknowns
                                      (<= 192, default 128; 0 = off) for a coarse-space
                                      correction over
                                      aggregates of consecutive device-order dofs, applied
                                      multiplicatively with the blocks (transposed in
                                      wb_solve_adjoint); changes iteration counts, not the
                                      solution */
    int assembly_strips;           /* 1 (default): the regular-cell assembly sweeps the
                                      logically structured patches of the mesh as strips of
                                      grid lines (every cell evaluated about once); 0: column
                                      tiles everywhere.  Same matrices either way */
    double ir_inner_lr_tol;        /* ir_inner_format = 2: a block is stored low rank when the
                                      range finder's estimate of its entry RMS error is at
                                      most this times the block's largest entry (1e-4) */
} wb_config;

typedef struct {
    int iterations;         /* GMRES iterations (FP32_IR: inner iterations, all sweeps) */
    int restarts;
    int ir_sweeps;          /* FP32_IR outer sweeps (0 for FP64) */
    double residual;        /* last preconditioned residual (convergence test value) */
    double true_rel_residual; /* ||b - A x|| / ||b|| of the constrained system, FP64 */
    double t_alpha_ms, t_rhs_ms, t_precond_ms, t_gmres_ms;
    int sweep_iterations[8];  /* FP32_IR: inner iterations of the first 8 sweeps */
    double sweep_residual[8]; /* FP32_IR: outer preconditioned residual before each sweep */
} wb_solve_stats;

typedef struct wb_ctx wb_ctx;

/* Defaults equal to the reference's parameter defaults (prm-files/default.prm:217-229)
 * and deal.II's GMRES AdditionalData(100). */
void wb_config_default(wb_config *cfg);

/* BEMProblem(ComputationalDomain&, BEMFMA&) (include/bem_problem.h:94) and
 * parse_parameters (:107).  Creates the context on cfg->device. */
wb_status wb_create(wb_ctx **ctx, const wb_config *cfg);
void wb_destroy(wb_ctx *ctx);
const char *wb_last_error(const wb_ctx *ctx);
const char *wb_version(void);

/* BEMProblem::reinit (source/bem_problem.cc:55-71) together with the mesh state it reads
 * from ComputationalDomain: support points (map_dofs_to_support_points,
 * :168-169), cells with their dofs in deal.II vertex order and direction flag, the
 * double_nodes_set (source/computational_domain.cc:257-307; CSR, each set contains the
 * dof itself; pass NULL to have it computed with the same 1e-8 coincidence rule), and
 * an optional per-cell material id (cell->material_id(), NULL allowed) used only to
 * group dofs in the device ordering.  Host buffers are copied. */
wb_status wb_set_mesh(wb_ctx *ctx, int n_dofs, int n_cells, const double *support_points,
                      const int32_t *cell_dofs, const int8_t *cell_orient,
                      const int32_t *cell_material, const int32_t *dn_ptr,
                      const int32_t *dn_idx);

/* Moves the support points of the current mesh (the reference's MappingQEulerian update
 * before each assemble_system, source/free_surface.cc:8744-8883): same cells, dofs and
 * double-node sets, so coincident dofs must stay coincident.  Keeps masks and constraints;
 * the next wb_assemble uses the new geometry. */
wb_status wb_set_support_points(wb_ctx *ctx, const double *support_points);

/* comp_dom.surface_nodes / other_nodes (source/numerical_towing_tank.cc:1799-1807,
 * source/free_surface.cc:8756-8772): 1 = Dirichlet (phi known) / Neumann. */
wb_status wb_set_masks(wb_ctx *ctx, const double *surface_nodes, const double *other_nodes);

/* The installed masks (as set by wb_set_masks or wb_prepare_bem_vectors); either pointer
 * may be NULL. */
wb_status wb_get_masks(wb_ctx *ctx, double *surface_nodes, double *other_nodes);

/* BEMProblem::assemble_system (source/bem_problem.cc:106-590): G (dirichlet_matrix) and
 * H (neumann_matrix) on the device, plus the solid-angle row sums. */
wb_status wb_assemble(wb_ctx *ctx);

/* No counterpart in the reference (BEMProblem keeps its FullMatrix members for its
 * lifetime).  Frees the device copies of G and H (and their FP32_IR lo halves and FP16
 * operator) so that a caller holding the context between a forward solve and its adjoint
 * can give the memory back; everything that reads the matrices returns WB_E_STATE until
 * the next wb_assemble, which allocates them again.  Mesh, masks, constraints and alpha
 * are kept. */
wb_status wb_release_matrices(wb_ctx *ctx);

/* Copies G and H back (row-major, n x n, caller dof order), for parity checks; either
 * pointer may be NULL. */
wb_status wb_get_matrices(wb_ctx *ctx, double *G, double *H);

/* The same for n_rows selected rows (caller numbering): G and H are n_rows x n_dofs,
 * row r holding matrix row rows[r].  For spot checks of large assemblies, where the
 * full matrices do not fit in host memory. */
wb_status wb_get_matrix_rows(wb_ctx *ctx, int n_rows, const int32_t *rows, double *G,
                             double *H);

/* No counterpart in the reference (its preconditioner is a band LU).  The coarse operator
 * Ac = P^T A P of the last preconditioner build (FP64, nc x nc row-major), A being the
 * block-Jacobi matrix (the vmult operator with identity rows for constrained dofs) in
 * the device dof order and P the indicator of the aggregates: device indices
 * [a agg, min(n, (a + 1) agg)) for a < nc.  For inspection; WB_E_STATE when no coarse
 * space is in use.  Ac may be NULL to query nc and agg. */
wb_status wb_get_coarse_matrix(wb_ctx *ctx, int *nc, int *agg, double *Ac);

/* BEMProblem::compute_alpha (source/bem_problem.cc:594-618): alpha = H * (-1). */
wb_status wb_compute_alpha(wb_ctx *ctx, double *alpha);

/* BEMProblem::vmult (source/bem_problem.cc:621-670), unconstrained. */
wb_status wb_vmult(wb_ctx *ctx, double *dst, const double *src);

/* BEMProblem::compute_rhs (source/bem_problem.cc:673-717), unconstrained. */
wb_status wb_compute_rhs(wb_ctx *ctx, double *dst, const double *src);

/* The ConstraintMatrix used by ConstrainedOperator (include/constrained_matrix.h:73-96),
 * as closed CSR lines: row line_rows[l] reads x_row - sum_p w[p] x[col[p]] and its
 * right-hand side is inhom[l].  Replaces any previous constraints. */
wb_status wb_set_constraints(wb_ctx *ctx, int n_lines, const int32_t *line_rows,
                             const int32_t *line_ptr, const int32_t *ent_col,
                             const double *ent_w, const double *inhom);

/* ComputationalDomain::compute_normals (source/computational_domain.cc:1525-1620) with
 * normalize = 1, NumericalTowingTank::compute_normals_at_nodes
 * (source/numerical_towing_tank.cc:3358-3433; the comp_dom.node_normals that
 * FreeSurface::prepare_bem_vectors reads) with normalize = 0: the L2 projection of the
 * cell normals onto the Q1 space at the current support points, solved on the device by
 * Jacobi-preconditioned CG on the mass matrix to a relative residual of 1e-14 (the
 * reference uses a direct solve).  normals: n x 3, caller order.  *cg_iterations (may be
 * NULL) receives the CG iteration count. */
wb_status wb_compute_normals(wb_ctx *ctx, int normalize, double *normals, int *cg_iterations);

/* FreeSurface::prepare_bem_vectors (source/free_surface.cc:8744-8883) for the current
 * mesh (its cell materials are required):
 *   - masks: surface_nodes = 1 on every dof of a cell whose material is in fs_ids, installed
 *     as by wb_set_masks;
 *   - on the dofs of the other cells, in cell order: wall cells (wall_ids)
 *     bem_bc = dphi_dn = -node_normals . wind, with node_normals the unnormalised
 *     projection of wb_compute_normals(normalize = 0) at the current support points and
 *     wind (n x 3) the wind function sampled at the support points; inflow cells
 *     (inflow_ids) inflow_dphi_dn (n, NULL reads as 0); other cells 0.  Free-surface
 *     entries of bem_bc / dphi_dn are left unchanged;
 *   - dofs with transom[i] != 0 (comp_dom.flags & transom_on_water; NULL = none) become
 *     Neumann with bem_bc = dphi_dn = the mean of prev_dphi_dn (the previous solution's
 *     bem_dphi_dn) over their coincident dofs. */
wb_status wb_prepare_bem_vectors(wb_ctx *ctx, int n_fs_ids, const int32_t *fs_ids,
                                 int n_wall_ids, const int32_t *wall_ids, int n_inflow_ids,
                                 const int32_t *inflow_ids, const double *wind,
                                 const double *inflow_dphi_dn, const uint8_t *transom,
                                 const double *prev_dphi_dn, double *bem_bc, double *dphi_dn);

/* BEMProblem::compute_surface_gradients (source/bem_problem.cc:1154-1290): the L2
 * projection of the surface gradient of phi = tmp_rhs .* surface_nodes (needs the masks),
 * solved as in wb_compute_normals.  grads: n x 3, caller order. */
wb_status wb_compute_surface_gradients(wb_ctx *ctx, const double *tmp_rhs, double *grads,
                                       int *cg_iterations);

/* FreeSurface::compute_potential_gradients (source/free_surface.cc:9271-9400): the L2
 * projection onto the Q1 space of the complete potential gradient n dphi_dn + surface
 * gradient of phi (full vectors, no masks), solved as in wb_compute_normals, then the mean
 * over every set of coincident dofs.  grads: n x 3, caller order. */
wb_status wb_compute_potential_gradients(wb_ctx *ctx, const double *phi, const double *dphi_dn,
                                         double *grads, int *cg_iterations);

/* BEMProblem::compute_constraints (source/bem_problem.cc:990-1105) for the current masks
 * and double nodes, from tmp_rhs.  The rules read node normals (n x 3,
 * ComputationalDomain::compute_normals) only where two Dirichlet dofs coincide and
 * surface gradients (n x 3, compute_surface_gradients(tmp_rhs)) only on sharp
 * Dirichlet-Dirichlet edges.  Pass them to use fixed values; pass NULL to have them
 * computed on the device as by wb_compute_normals / wb_compute_surface_gradients at the
 * current support points (and, when the constraints are rebuilt by wb_solve_system for a
 * new tmp_rhs, again from that tmp_rhs), as the reference does at the start of
 * compute_constraints.  Installs the result as with wb_set_constraints and returns the
 * number of lines in *n_lines (may be NULL). */
wb_status wb_compute_constraints(wb_ctx *ctx, const double *tmp_rhs, const double *node_normals,
                                 const double *node_surface_gradients, int *n_lines);

/* ConstrainedOperator<Vector,BEMProblem>::vmult: dst = A src with constrained rows
 * replaced. */
wb_status wb_constrained_vmult(wb_ctx *ctx, double *dst, const double *src);

/* No counterpart in the reference (it solves in FP64 only).  The operator of the inner
 * Krylov solves of an FP32_IR context (per ir_inner_format: the FP32 hi operator, its
 * row-scaled FP16 copy or its compressed copy) applied to src in FP32 vectors, constrained
 * rows as in wb_constrained_vmult; src and dst in caller numbering.  For inspection of the
 * inner operator's error; WB_E_UNSUPPORTED in FP64 contexts. */
wb_status wb_inner_vmult(wb_ctx *ctx, double *dst, const double *src);

/* No counterpart in the reference.  The compressed inner operator (ir_inner_format 2) of
 * the current assembly and masks, built if needed: info[0] clusters, [1] dense blocks,
 * [2] low-rank blocks, [3] zero blocks, [4] sum of the low-rank blocks' ranks, [5] bytes
 * of factors and dense blocks one product reads.  WB_E_UNSUPPORTED for other formats. */
wb_status wb_inner_operator_info(wb_ctx *ctx, int64_t *info);

/* No counterpart in the reference.  wb_constrained_vmult with src and dst device pointers
 * (caller numbering, n doubles each) that allocates no device memory, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops): its work space is
 * a fixed static device array per GPU (two vectors of at most 262144 doubles), so
 * n rounded up to 64 must not exceed 262144 (WB_E_ARG otherwise); the context must be
 * assembled, with its masks and constraints installed; pure-Neumann operators are
 * rejected (WB_E_UNSUPPORTED).  Returns after the product is complete.  Calls that
 * overlap on one GPU share the work space and race. */
wb_status wb_constrained_vmult_noalloc(wb_ctx *ctx, const double *src_dev, double *dst_dev);

/* BEMProblem::solve_system (source/bem_problem.cc:821-895): alpha, rhs from tmp_rhs,
 * distribute_rhs, preconditioned GMRES (sol = 0 start), then phi(i) = sol(i) where
 * surface_nodes(i) == 0 else dphi_dn(i) = sol(i).  Uses the installed constraints.
 * On WB_E_NOCONV phi/dphi_dn are left unchanged, as the reference throws before the
 * scatter.  stats may be NULL. */
wb_status wb_solve_system(wb_ctx *ctx, const double *tmp_rhs, double *phi, double *dphi_dn,
                          wb_solve_stats *stats);

/* Same with device pointers (n doubles each, internal stream ordering respected). */
wb_status wb_solve_system_dev(wb_ctx *ctx, const double *tmp_rhs_dev, double *phi_dev,
                              double *dphi_dn_dev, wb_solve_stats *stats);

/* BEMProblem::solve (source/bem_problem.cc:969-987), Direct branch: assemble_system +
 * solve_system. */
wb_status wb_solve(wb_ctx *ctx, const double *tmp_rhs, double *phi, double *dphi_dn,
                   wb_solve_stats *stats);

/* BEMProblem::residual (source/bem_problem.cc:903-961), Direct branch, with the
 * installed constraints. */
wb_status wb_residual(wb_ctx *ctx, double *res, const double *phi, const double *dphi_dn);

/* ---- adjoint (no reference counterpart: the reference differentiates its residual with
 * Sacado, source/free_surface.cc:4918-5053, 6099-6109) ------------------------------ */

/* Solves A_c^T lambda = rhs with the transposed ConstrainedOperator of the installed
 * constraints and masks (the operator of solve_system), preconditioned by the transposed
 * block Jacobi (FP64: with the transposed coarse space), with the same GMRES settings.
 * FP32_IR contexts refine in FP64 with the transposed hi + lo operator around an inner
 * FP32 GMRES on the transposed hi part (the FP16 and compressed inner operators hold the
 * forward selection only).  rhs and lambda in the caller's dof numbering.  All-Neumann
 * masks return WB_E_UNSUPPORTED; WB_E_NOCONV leaves lambda unchanged. */
wb_status wb_solve_adjoint(wb_ctx *ctx, const double *rhs, double *lambda,
                           wb_solve_stats *stats);

/* Backward of assemble_system with respect to the support points: grad (n x 3, caller
 * numbering) = sum_i w_i dR_i/dx_d for every support point x_d, where
 *   R_i = (alpha_i phi_i + sum_j H_ij phi_j) - sum_j G_ij dphi_dn_j
 * is the BEM row of vmult/compute_rhs on the full vectors (alpha = -H 1), with the
 * quadrature of the current assembly (regular and singular cells).  A point moves as the
 * collocation node of its row and as the vertex of every cell that uses its dof; callers
 * moving coincident dofs together add their rows.  With free_rows_only the rows of the
 * installed constraint lines get weight 0 (their residual does not depend on geometry).
 * Needs wb_assemble at the current geometry. */
wb_status wb_residual_vjp(wb_ctx *ctx, const double *w, const double *phi,
                          const double *dphi_dn, int free_rows_only, double *grad);

/* Tangent of the same residual: dres (n, caller numbering) with
 * dres_i = sum_d dR_i/dx_d . dx_d for a direction dx (n x 3) of the support points; rows of
 * the installed constraint lines are 0 with free_rows_only.  Needs wb_assemble at the
 * current geometry. */
wb_status wb_residual_jvp(wb_ctx *ctx, const double *dx, const double *phi,
                          const double *dphi_dn, int free_rows_only, double *dres);

/* Solves A_c sol = rhs with the operator, preconditioner and solver of solve_system for a
 * given right-hand side (no boundary-data mapping, no constraint inhomogeneities): the
 * linear solve of a tangent, e.g. du = -A_c^-1 (dR/dx dx) for a Newton JVP.  sol holds the
 * unknowns (phi on Neumann dofs, dphi_dn on Dirichlet dofs), caller numbering. */
wb_status wb_solve_rhs(wb_ctx *ctx, const double *rhs, double *sol, wb_solve_stats *stats);

/* Multi right-hand side form of solve_system (no reference counterpart: the reference
 * calls solve_system once per time step, source/bem_problem.cc:1013-1108): k
 * boundary-data vectors (column j at tmp_rhs + j*n), solved in batches of up to
 * WB_MULTI_MAX with one batched (MFMA) matrix product per GMRES step and per-column
 * GMRES state (each column converges, restarts and fails on its own, as k separate
 * solve_system calls would).  Constraint inhomogeneities: with automatic constraints
 * they are rebuilt from each column of tmp_rhs; otherwise taken from inhom_k (n_lines x k,
 * column-major; NULL = the installed ones for every column).  All-Neumann masks
 * return WB_E_UNSUPPORTED. */
#define WB_MULTI_MAX 16
wb_status wb_solve_system_multi(wb_ctx *ctx, int k, const double *tmp_rhs, double *phi,
                                double *dphi_dn, const double *inhom_k,
                                wb_solve_stats *stats);

/* ---- kernel timing (HIP events on the context stream around each hot launch) ------- */
typedef enum {
    WB_K_ASSEMBLE = 0,  /* regular-cell G/H tile kernel */
    WB_K_SINGULAR = 1,  /* singular-cell kernel */
    WB_K_GEMV64 = 2,    /* column-selected GEMV, FP64 matrices */
    WB_K_GEMV32 = 3,    /* column-selected GEMV, FP32 (hi) matrices, inner solver */
    WB_K_GEMVHL = 4,    /* column-selected GEMV, FP32 hi + bf16 lo, FP64 accumulation */
    WB_K_GEMV16 = 5,    /* GEMV of the row-scaled FP16 operator, inner solver */
    WB_K_GEMVC = 6,     /* product of the compressed operator (ir_inner_format 2) */
    WB_K_INNER = 7,     /* build of the inner operator (FP16 copy or compressed copy) */
    WB_K_COUNT = 8
} wb_kernel_kind;

typedef struct {
    int launches[WB_K_COUNT];   /* timed launches */
    double total_ms[WB_K_COUNT];
    long long rows[WB_K_COUNT]; /* GEMV: matrix rows read, summed over timed launches */
    int calls[WB_K_COUNT];      /* all launches.  GEMV64, GEMV32, GEMV16 and GEMVC (one per
                                   GMRES step) are timed on one launch in 8: each event pair
                                   idles the GPU for about 11 us, against 70 us for the
                                   compressed product; the others on every launch */
} wb_kernel_stats;

/* enable != 0 starts (and resets) event timing of the hot kernels; 0 stops it. */
wb_status wb_set_profiling(wb_ctx *ctx, int enable);
/* Synchronizes, folds all recorded events into *out (accumulated since enabling). */
wb_status wb_get_kernel_stats(wb_ctx *ctx, wb_kernel_stats *out);

/* ---- host-only utilities (no device needed) ---------------------------------------- */

/* The quadrature tables the device path uses: QGauss<2>(q) (q^2 points, xi fastest,
 * reg_pts[2q+{0,1}]) and QGaussOneOverR<2>(s, vertex v, factor_out=true) for v = 0..3
 * (2 s^2 points each, vertex-major: sing_pts[2(v*2s^2+k)+{0,1}]); as requested at
 * source/bem_problem.cc:116-120,125-137.  Any pointer may be NULL. */
int wb_host_quadrature(int q_order, int s_order, double *reg_pts, double *reg_w,
                       double *sing_pts, double *sing_w);

/* The device dof ordering (perm[dof] = device index), the double_nodes_set computed with
 * the 1e-8 coincidence rule of source/computational_domain.cc:257-307 (CSR), and the
 * sizes of the assembly plan.  Any output pointer may be NULL. */
int wb_host_mesh_plan(int n_dofs, int n_cells, const double *support_points,
                      const int32_t *cell_dofs, const int32_t *cell_material, int32_t *perm,
                      int32_t *dn_ptr, int32_t *dn_idx, int dn_cap, int32_t *n_tiles,
                      int32_t *n_tile_cells, int32_t *n_singular_pairs);

/* No counterpart in the reference (its assembly loops over cells serially).  The strip
 * schedule of the regular-cell assembly for inspection: sizes[0..4] = tasks, gaps, cell
 * evaluations, strip dofs, column tiles.  With non-NULL buffers (of at least those sizes):
 * tasks (2 ints each: first gap, gaps), gaps (4 ints: cell mask, first device column of
 * the line completed by the gap, its position mask, 0), gap_cells (9 ints per gap: cell at
 * each position or -1), cell_int_s (4 ints per cell: device vertex indices in the
 * canonical order) and perm (dof -> device index).  strips = 0 plans column tiles only. */
int wb_host_strip_plan(int n_dofs, int n_cells, const double *support_points,
                       const int32_t *cell_dofs, const int32_t *cell_material, int strips,
                       int32_t *sizes, int32_t *tasks, int32_t *gaps, int32_t *gap_cells,
                       int32_t *cell_int_s, int32_t *perm);

/* No counterpart in the reference.  The cluster plan of the compressed inner operator
 * (ir_inner_format 2) for inspection: *n_clusters always; with non-NULL buffers (of at
 * least that many clusters) cl_size, cl_idx (256 device indices per cluster, -1 past its
 * size), near (n_clusters^2 flags: the block is dense by geometry) and perm (dof ->
 * device index). */
int wb_host_cluster_plan(int n_dofs, int n_cells, const double *support_points,
                         const int32_t *cell_dofs, const int32_t *cell_material,
                         int32_t *n_clusters, int32_t *cl_size, int32_t *cl_idx, uint8_t *near,
                         int32_t *perm);

/* No counterpart in the reference.  The work lists of the compressed operator's product
 * for inspection, from block ranks (n_clusters^2 bytes, block (I, J) at I n_clusters + J:
 * 0 zero, 255 dense, else the rank) and cluster sizes: sizes[0..4] = chunks per row cluster,
 * items, columns per dense item, item shift, low-rank kind.  With non-NULL buffers (of
 * n_clusters chunks + 1 and items entries), item_ptr and items: the items of row cluster I
 * and chunk k are items[item_ptr[I chunks + k] .. item_ptr[I chunks + k + 1]), each
 * (J << shift) | kind, kind = low-rank kind for block (I, J) or q for its dense columns
 * q cols .. (q + 1) cols - 1; low-rank items first in each chunk. */
int wb_host_hm_work_items(int n_clusters, const uint8_t *rank, const int32_t *cl_size,
                          int32_t *sizes, int32_t *item_ptr, int32_t *items);

/* BEMProblem::compute_constraints (source/bem_problem.cc:990-1105) rules, closed as by
 * ConstraintMatrix::close; returns the number of lines, or minus a wb_status. */
int wb_host_compute_constraints(int n, const int32_t *dn_ptr, const int32_t *dn_idx,
                                const double *surface, const double *node_normals,
                                const double *node_surface_gradients, const double *tmp_rhs,
                                int cap_lines, int cap_nnz, int32_t *line_rows, int32_t *line_ptr,
                                int32_t *ent_col, double *ent_w, double *inhom);

/* Device memory used by the context in bytes. */
size_t wb_device_bytes(const wb_ctx *ctx);

/* Synchronizes the context stream. */
wb_status wb_synchronize(wb_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
