// wb_internal.h -- internal state and kernel launchers of libwavebem_bem.
//
// Device layout (all in the internal dof order chosen by wb_setup.cpp: dofs grouped by
// cell material, then Morton order of the support point, so that consecutive internal
// indices are spatial neighbours and patches with one boundary-condition type are
// contiguous):
//   G, H        n x ld row-major (ld = n rounded up to 64).  FP64 mode: double.
//               FP32_IR mode: hi = float, lo = bf16 bits of (A - hi) (uint16).
//   cellq       [n_cells][NQ][8] doubles: quadrature point y(3), n*JxW(3), JxW, pad.
//   column tiles of TILE_COLS consecutive internal columns; for each tile the list of
//               cells having a vertex dof in it, with the 4 tile-local columns packed
//               as int8 (-1 = vertex dof outside the tile).
//   singular pairs (row, cell, vertex) CSR by row.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/wavebem_bem.h"

namespace wb {

constexpr int TILE_ROWS = 64;   // rows per assembly workgroup (one wavefront, lane = row)
constexpr int TILE_COLS = 16;   // columns per assembly register tile
constexpr int TILES_PER_WG = 8; // column tiles swept by one workgroup
constexpr int MAX_NQ1 = 6;      // supported regular QGauss orders 1..6
constexpr int MAX_SING = 8;     // supported singular QGaussOneOverR orders 1..8
constexpr double INV4PI = 0.07957747154594767;  // 1/(4 pi)

struct PhiTable {
    double phi[MAX_NQ1 * MAX_NQ1 * 4];  // phi[q*4+k], regular rule shape values
};

struct Quadrature {
    int q_order = 0, s_order = 0;
    std::vector<double> reg_pts, reg_w;    // nq x 2, nq
    std::vector<double> sing_pts, sing_w;  // 4 x ns x 2, 4 x ns (vertex-major)
};

// Host quadrature tables (deal.II QGauss<2> and QGaussOneOverR<2> semantics).
void build_quadrature(int q_order, int s_order, Quadrature &q);
void gauss_legendre_01(int n, double *x, double *w);

// Strip schedule of the regular-cell assembly on logically structured patches.  A line
// is at most STRIP_W dofs of one grid row, at consecutive positions; a strip is the lines
// of one band of positions over consecutive grid rows ("levels"); a task is a run of
// levels of one strip, swept by one wavefront per 64 collocation rows.  The cells of a
// task lie between consecutive levels, so two lines at a time hold accumulators: the
// ring of 2 STRIP_W slots, line l in half l & 1.
constexpr int STRIP_W = 8;
#ifndef WB_STRIP_TASK_ENTRIES
#define WB_STRIP_TASK_ENTRIES 640
#endif
constexpr int STRIP_TASK_ENTRIES = WB_STRIP_TASK_ENTRIES;  // cells per task, about

struct MeshPlan {
    int n = 0, nc = 0;
    size_t ld = 0;
    std::vector<int> perm, iperm;          // dof -> internal, internal -> dof
    std::vector<int> cls;                  // per internal index: double-node class id
    std::vector<int> cell_cls;             // nc x 4 class ids of cell vertices
    std::vector<int> cell_int;             // nc x 4 internal vertex indices
    std::vector<int> tile_ptr, tile_cell, tile_cols, tile_start;
    std::vector<int> sing_ptr, sing_cell, sing_vtx;  // by internal row
    int n_tiles = 0;
    int tc = TILE_COLS;                    // columns per tile
    std::vector<int> dn_ptr, dn_idx;       // caller-numbered double node sets
    // strips (structured patches): tasks {first gap, gaps}; gap t (task-local -1 .. L-1,
    // L lines) = the cells between task lines t and t + 1 (STRIP_W + 1 positions: cell k
    // touches line positions k - 1 and k) as {cell mask, first column of line t, position
    // mask of line t (0 for t = -1), 0}, and their cell ids st_gcell (STRIP_W + 1 per gap,
    // -1 = none).  Cells of structured patches have their vertices in the canonical order
    // cell_int_s (position fastest, then level), canon_sign = the orientation change of
    // that reordering; elsewhere cell_int_s = cell_int, canon_sign = 1.
    int n_tasks = 0;
    long strip_cells = 0;                  // cell evaluations of all tasks
    std::vector<int> st_task, st_gap, st_gcell;
    std::vector<int> cell_int_s;
    std::vector<int8_t> canon_sign;
    int strip_dofs = 0;                    // dofs covered by strips (device order [0, this))
    std::vector<int> run_start;            // device index where each strip begins
};

std::string plan_mesh(int n, int nc, const double *support, const int32_t *cell_dofs,
                      const int32_t *cell_material, const int32_t *dn_ptr, const int32_t *dn_idx,
                      int tile_cols, int strips, MeshPlan &plan);

// ---------------------------------------------------------------- compressed inner operator
// The inner solves of FP32 iterative refinement may run on a compressed copy of the vmult
// operator.  Its rows and columns are split into clusters of at most HM_CL dofs, each the
// union of at most HM_SEGS segments (runs of at most HM_SEG consecutive device indices
// inside one strip, or consecutive column tiles), grouped by geometric bisection; block
// (I, J) of the cluster partition is stored dense (FP16) when the clusters' bounding boxes
// touch, and otherwise as Q B with rank <= HM_RMAX when a randomized range finder (HM_SK
// sign vectors) meets the tolerance, else dense too.
constexpr int HM_SEG = 64, HM_SEGS = 4, HM_CL = HM_SEG * HM_SEGS, HM_SK = 16, HM_RMAX = 12;
constexpr int HM_DENSE = 255;
// product work items: (J << HM_ITEM_SHIFT) | kind, kind q < HM_CL / HM_DCOLS for columns
// HM_DCOLS q .. HM_DCOLS (q + 1) - 1 of a dense block, HM_ITEM_LOWRANK for a low-rank block
#ifndef WB_HM_DCOLS
#define WB_HM_DCOLS 16
#endif
constexpr int HM_DCOLS = WB_HM_DCOLS;
constexpr int HM_ITEM_SHIFT = 5, HM_ITEM_MASK = 31, HM_ITEM_LOWRANK = 31;
#ifndef WB_HM_SKETCH_PARTS
#define WB_HM_SKETCH_PARTS 2
#endif
// wavefronts per 16-row block of the sketch (cluster ranges)
constexpr int HM_SKETCH_PARTS = WB_HM_SKETCH_PARTS;
struct HmatPlan {
    int ncl = 0;
    std::vector<int> cl_idx;    // ncl x HM_CL device indices (-1 past the cluster's size)
    std::vector<int> cl_size;   // ncl
    std::vector<int> cl_seg;    // ncl x HM_SEGS x {first device index, length} (0, 0 unused)
                                // (on the device with the column kind of the masks, int4)
    std::vector<uint8_t> near;  // ncl x ncl: 1 = dense by geometry
};
void plan_hmat(const MeshPlan &P, const double *X_internal, HmatPlan &H);
// Build (wb_hmat.hip): Y (ncl x n x HM_SK) and Mx (ncl x n) are work space; rank and
// fscale (ncl^2), Q16 / B16 (ncl^2 x HM_RMAX x HM_CL halfs) and D16 (ncl^2 x HM_CL x HM_CL
// halfs) hold the operator.  With S (agg >= HM_SEG; HM_SKETCH_PARTS x nagg x ldS doubles of
// work space), the build also writes S[a ldS + i] = the sum of row i of the FP32 operator
// over the columns of aggregate a (a < nagg), the coarse space's A P.  Product: y = A~ x + diag .* xd (FP32 vectors, all n rows); xc
// (ncl x HM_CL, x in cluster order; with xc_ready already holding it) and ypart
// (hm_chunks(ncl) x ldp) floats are work space.  With line (constraint line per row, -1
// free; cptr/ccols/cw the lines): constrained rows get x_r - sum w x_col instead.  With
// Cpart (2 ceil(n / 64) doubles): per 64-row block, the sums of y over the two aggregates
// of agg rows starting at the block's first (for launch_bj_coarse).
int hm_chunks(int ncl);
// build, in two launches with the host between them: launch_hm_rank (sketch, aggregate
// sums, rank test, Q) writes rank[]; the caller lays out the dense blocks (doff[b]: the
// block's slot in D16, HM_CL x HM_CL halfs each; -1 if not dense) and launch_hm_fill
// writes B and the dense blocks
void launch_hm_rank(const float *G, const float *H, size_t ld, int n, const uint8_t *gmask,
                    const int *cl_idx, const int *cl_size, const int *cl_seg, int ncl,
                    const uint8_t *near, float tol, float *Y, float *Mx, uint8_t *rank,
                    float *fscale, void *Q16, int agg, int nagg, double *S, size_t ldS,
                    hipStream_t s);
void launch_hm_fill(const float *G, const float *H, size_t ld, const uint8_t *gmask,
                    const int *cl_idx, const int *cl_size, int ncl, const uint8_t *rank,
                    float *fscale, const void *Q16, void *B16, void *D16, const int *doff,
                    hipStream_t s);
// Product work lists from the ranks (host copies): for row cluster I and chunk k <
// hm_chunks(ncl), items[item_ptr[I chunks + k] .. item_ptr[I chunks + k + 1]) are the
// work items of row cluster I (low-rank blocks and column parts of dense blocks), spread
// over its chunks by weight (bytes read), low-rank ones first
void hm_plan_items(int ncl, const uint8_t *rank, const int *cl_size, std::vector<int> &item_ptr,
                   std::vector<int> &items);
void launch_hm_gemv(const int *cl_idx, const int *cl_size, int ncl, const uint8_t *rank,
                    const float *fscale, const void *Q16, const void *B16, const void *D16,
                    const int *doff, const int *item_ptr, const int *items, int n,
                    const float *x, const double *diag, const float *xd, float *xc, float *ypart,
                    size_t ldp, float *y, hipStream_t s, bool xc_ready = false,
                    const int *line = nullptr, const int *cptr = nullptr,
                    const int *ccols = nullptr, const double *cw = nullptr, int agg = 64,
                    double *Cpart = nullptr);

// The compressed product for k vectors (columns of X and Y at strides ldx / ldy, FP32),
// read once per group of up to 8 vectors: Y = A~ X + diag .* X, with the constraint lines
// (line / cptr / ccols / cw as launch_hm_gemv, or line = nullptr).  Work space: xcm and
// ypart of the sizes hm_multi_work_floats returns (floats).
size_t hm_multi_work_floats(int ncl, size_t ldp, size_t *xc_floats);
void launch_hm_gemv_multi(const int *cl_idx, const int *cl_size, int ncl, const uint8_t *rank,
                          const float *fscale, const void *Q16, const void *B16,
                          const void *D16, const int *doff, const int *item_ptr,
                          const int *items, int n, int k, const float *X, size_t ldx,
                          const double *diag, float *xcm, float *ypart, size_t ldp, float *Y,
                          size_t ldy, const int *line, const int *cptr, const int *ccols,
                          const double *cw, hipStream_t s);

// Work space of the entry points that allocate nothing (wb_noalloc.hip): a static device
// array of 2 NOALLOC_MAX_N doubles per GPU.  noalloc_setup looks its address up on a device
// (once; called when a context is created there); noalloc_scratch returns it (nullptr
// before setup).
constexpr int NOALLOC_MAX_N = 262144;
hipError_t noalloc_setup(int device);
double *noalloc_scratch(int device);

// Constraint rules of BEMProblem::compute_constraints, caller numbering, closed.
struct ConstraintSet {
    std::vector<int> rows, ptr, cols;
    std::vector<double> w, inhom;
};
std::string build_constraints(int n, const int *dn_ptr, const int *dn_idx, const double *surface,
                              const double *normals, const double *grads, const double *tmp_rhs,
                              ConstraintSet &out);

// ---------------------------------------------------------------- kernel launchers
// Records of n_slots cells: slot s holds cell s, or cell slot_cell[s] when slot_cell is
// given (slots with -1 are left untouched).
void launch_cell_geometry(const double *support_int, const int *cell_int, const int8_t *orient,
                          int n_slots, int nq, const double *qpts, const double *qw,
                          double *cellq, hipStream_t s, const int *slot_cell = nullptr);
void launch_assemble_tiles(int precision, int q_order, int tc, const double *xrow,
                           const double *cellq, const int *tile_ptr, const int *tile_cell,
                           const int *tile_cols, const int *tile_start, int n, int n_tiles,
                           size_t ld, const PhiTable &phi, void *G, void *H, void *Glo,
                           void *Hlo, double *hsum_part, hipStream_t s);
// Strip tasks (MeshPlan::st_*): recs holds the cell records of every gap position
// (gap g, position k at record g * (STRIP_W + 1) + k), built from cell_int_s;
// hsum_part rows [0, n_tasks) receive the tasks' row sums of H.
void launch_assemble_strips(int precision, int q_order, const double *xrow, const double *recs,
                            const int *tasks, const int *gaps, int n, int n_tasks, size_t ld,
                            const PhiTable &phi, void *G, void *H, void *Glo, void *Hlo,
                            double *hsum_part, hipStream_t s);
// The regular kernels add every cell with the regular rule, singular ones included; this
// kernel replaces, per singular (row, cell) pair, that regular part (recomputed from the
// records cellq, nq points at the reference points qpts) by the QGaussOneOverR integral.
void launch_assemble_singular(int precision, const double *support_int, const int *cell_int,
                              const int8_t *orient, const int *sing_ptr, const int *sing_cell,
                              const int *sing_vtx, int n, int ns, const double *sing_rule,
                              const double *cellq, int nq, const double *qpts,
                              size_t ld, void *G, void *H, void *Glo, void *Hlo,
                              double *hsum_sing, hipStream_t s);
// L2 projections onto the Q1 space (wb_project.hip); vectors [n][6]: normal x,y,z and
// surface gradient x,y,z.  state holds rho[2][6], |b|^2[6], |r|^2[6].
int proj_blocks(int n);
void launch_proj_setup(const double *X, const int *cell_int, const int8_t *orient,
                       const double *u, const double *psi, int nc, int nq, const double *qpts,
                       const double *qw, const int *inc_ptr, const int *inc_item, int n,
                       double *emass, double *erhs, double *B, double *dinv, hipStream_t s);
void launch_proj_cg_start(const double *B, const double *dinv, int n, double *Xs, double *R,
                          double *Z, double *P, double *part2, double *state, hipStream_t s);
void launch_proj_cg_iter(const int *inc_ptr, const int *inc_item, const int *cell_int,
                         const double *emass, const double *dinv, int n, int iter, double *Xs,
                         double *R, double *Z, double *P, double *Q, double *part,
                         double *part2, double *state, hipStream_t s);
void launch_alpha_reduce(const double *hsum_part, int n_parts, const double *hsum_sing, int n,
                         double *alpha, hipStream_t s);

// y[r] = sum_j M_sel(j)[r,j] * sign_sel(j) * x[j] + diag[r] * xd[r] for r in rows (or all)
// sel bit per column pair in pairflag (0: G,G 1: H,H 2: G,H 3: H,G), flipped by `flip`.
// mode: 0 = FP64 matrices, 1 = hi only (FP32 accumulate), 2 = hi+lo (FP64 accumulate)
// segs: n_segs x int4 {first 4-column group, end group, kind 0=G 1=H 2=mixed, 0}
void launch_gemv_sel(int mode, const void *G, const void *H, const void *Glo, const void *Hlo,
                     size_t ld, const int *segs, int n_segs, const uint8_t *gmask, int flip,
                     double s_sel, double s_unsel, const void *x, const double *diag,
                     const void *xd, const int *rows, int n_rows, void *y, hipStream_t s);
// k <= 16 vectors at once (column v at X + v*ldx, Xd + v*ldx, Y + v*ldy) on the matrix
// cores; the columns are split into multi_chunks(mode, n_rows, n_groups) chunks whose
// partials (part: chunks * k * n_rows doubles) are added in order.  Mode 3 (FP16 operator)
// takes G = A16, H = its row scales and Glo = a workspace of 256 + 4 k ld bytes (the vector
// scales, then the FP16 hi and lo halves of the vectors).
int multi_chunks(int mode, int n_rows, int n_groups);
// FP16 inner operator (wb_gemv.hip): A16 (n x ld halfs, row scales rs) from the FP32
// column-selected vmult operator, and y = diag(rs) A16 x + diag .* xd over the listed rows.
// With S non-null (agg a multiple of 4) the build may also write S[a ldS + i] = sum of row
// i of diag(rs) A16 over the columns of aggregate a (a < nc); it returns whether it did.
bool launch_build_h16(const float *G, const float *H, size_t ld, int n, const int *segs,
                      int n_segs, const uint8_t *gmask, void *A16, float *rs, int agg, int nc,
                      double *S, size_t ldS, hipStream_t s);
void launch_gemv_h16(const void *A16, const float *rs, size_t ld, int n, const float *x,
                     const double *diag, const float *xd, const int *rows, int n_rows, float *y,
                     hipStream_t s);
// Blocks of `threads` threads of `kernel` the current device holds at once (occupancy
// query times the CU count), cached per kernel, block size and device; 0 when the runtime
// cannot say (callers then keep a fixed grid shape).
int resident_blocks(const void *kernel, int threads);
void launch_gemv_sel_multi(int mode, const void *G, const void *H, const void *Glo,
                           const void *Hlo, size_t ld, const int *segs, int n_segs,
                           const uint8_t *gmask, int flip, double s_sel, double s_unsel, int k,
                           const void *X, size_t ldx, const double *diag, const void *Xd,
                           const int *rows, int n_rows, void *Y, size_t ldy, double *part,
                           int n_groups, hipStream_t s);
// constrained rows: y[row] = x[row] - sum w x[col]  (mode_inhom=0)  or y[row] = inhom (1)
// op = alpha .* (1 - s), rhs = -alpha .* s (n entries, s a 0/1 mask)
void launch_diag_from_mask(int n, const double *alpha, const double *s, double *op, double *rhs,
                           hipStream_t st);
// a[i] = v[i] where mask[i] == 0, b[i] = v[i] elsewhere (n entries)
void launch_scatter_mask(int n, const uint8_t *mask, const double *v, double *a, double *b,
                         hipStream_t s);
void launch_constraint_rows(int is_float, int n_lines, const int *rows, const int *ptr,
                            const int *cols, const double *w, const double *inhom, int mode_inhom,
                            const void *x, void *y, hipStream_t s);
void launch_block_jacobi_build(int n, int bs, const void *G, const void *H, const void *Glo,
                               const void *Hlo, int precision, size_t ld, const double *alpha,
                               const double *surface,
                               const uint8_t *constrained, double *binv, float *binv32,
                               hipStream_t s);
// With C (CS_MAX doubles per vector, see launch_coarse_pre): y_i += (Aci C[v])[i / agg],
// the coarse-space part (Aci: nc x nc row-major, FP32 for is_float; with trans, Aci^T C).
void launch_block_jacobi_apply(int is_float, int n, int bs, const void *binv, const void *x,
                               void *y, int k, size_t ldx, hipStream_t s, int trans = 0,
                               const double *C = nullptr, const void *Aci = nullptr, int nc = 0,
                               int agg = 1);

// ---------------------------------------------------------------- coarse space (wb_coarse.hip)
// Aggregates of agg consecutive device-order dofs (agg a multiple of 8, at least 64),
// nc <= CS_MAX of them (the inverse is eliminated in FP64 up to CS_MAX64, in FP32 above).
// Build: AP (nc x ldap, [a][j] = sum over aggregate a of the block-Jacobi matrix row j;
// FP64 matrices, the FP32 hi part, with A16 != nullptr the row-scaled FP16 copy, or with
// S != nullptr the FP16 copy's aggregate sums written by launch_build_h16),
// Ac = P^T AP, Aci = Ac^-1 (nc x nc row-major),
// WT ([b][j] = (AP Aci)[j][b]), each in FP64 and FP32; status[0] = 0 when the inverse is
// usable.  Apply (k vectors, column stride ldx): C[v] = P^T x_v (CS_MAX doubles per vector)
// and Tout_v = x_v - W C[v]; the block-Jacobi apply then adds P Aci C[v].
constexpr int CS_MAX = 192, CS_MAX64 = 128;
// C = P^T x alone (k vectors)
void launch_coarse_restrict(int is_float, int n, int agg, int nc, const void *X, size_t ldx,
                            double *C, int k, hipStream_t s);
// One vector: y = B^-1 (x - W C) + P Aci C in one kernel (x and y may alias); usable when
// bj_coarse_fits(bs, agg, nc)
// bj_coarse_fits(bs, agg, nc); with Cpart (see launch_hm_gemv), C = P^T x is summed from
// those block partials instead of read from C
bool bj_coarse_fits(int bs, int agg, int nc);
void launch_bj_coarse(int is_float, int n, int bs, const void *binv, const void *x, void *y,
                      const void *WT, size_t ldw, const double *C, const void *Aci, int nc,
                      int agg, hipStream_t s, const double *Cpart = nullptr);
size_t coarse_invert_lds_bytes(int nc);
void launch_coarse_build(int n, int agg, int nc, const void *G, const void *H, int precision,
                         size_t ld, const void *A16, const float *rs16, const double *alpha,
                         const uint8_t *gmask, const uint8_t *constrained, double *AP,
                         size_t ldap, const double *S, double *Ac, double *Aci,
                         float *Aci32, double *WT, float *WT32, int *status, hipStream_t s);
void launch_coarse_pre(int is_float, int n, int agg, int nc, const void *X, size_t ldx,
                       const void *WT, size_t ldw, double *C, void *Tout, int k, hipStream_t s);
// The transposed system (FP64): PTA = P^T A (nc x ldap, [a][i]) and WTT ([b][i] =
// (A^T P Ac^-T)[i][b]) from the same aggregates and Aci; the block-Jacobi apply then adds
// Ac^-T C (launch_block_jacobi_apply with trans = 1).
void launch_coarse_build_t(int n, int agg, int nc, const double *G, const double *H, size_t ld,
                           const double *alpha, const uint8_t *gmask, const uint8_t *constrained,
                           const double *Aci, double *PTA, size_t ldap, double *WTT,
                           hipStream_t s);

// ---------------------------------------------------------------- adjoint (wb_adjoint.hip)
// y = A_sel^T x over the unconstrained rows + diag .* x on those rows; mode 0: FP64
// matrices and vectors, 1: FP32 hi part and FP32 vectors, 2: hi + bf16 lo and FP64 vectors
// (FP64 sums in every mode); part: gemvT_chunks(n) x ld doubles
int gemvT_chunks(int n);
void launch_gemvT(int mode, const void *G, const void *H, const void *Glo, const void *Hlo,
                  size_t ld, int n, const uint8_t *gmask, int flip, double s_sel,
                  double s_unsel, const void *x, const double *diag,
                  const uint8_t *constrained, double *part, void *y, hipStream_t s);
void launch_constraint_T(int is_float, int n_targets, const int *tgt, const int *ptr,
                         const int *src, const double *coef, const void *x, void *y,
                         hipStream_t s);
// backward of the assembly; buffer sizes: cstate nc*nq*2, part1 vjp_row_chunks(nc)*n*3,
// part2 vjp_point_chunks(n)*nc*nq*7, xgrad n*3, vbar nc*12, xs n*3, spart n_sing*12
int vjp_row_chunks(int nc);
int vjp_point_chunks(int n);
void launch_residual_vjp(int q_order, int n, int nc, const double *X, const int *row_cls,
                         const double *wrow, const double *phi, const double *psi,
                         const double *rowpack, const double *cellq, const int *cell_int,
                         const int *cell_cls, const int8_t *orient, const double *qpts,
                         const double *qw, const int *sing_ptr, const int *sing_cell,
                         const int *sing_vtx, int ns, const double *srule, double *cstate,
                         double *part1, double *part2, double *xgrad, double *vbar,
                         double *xs, double *spart, hipStream_t s);

// Krylov helpers (is_float selects float vectors).  batch instances on blockIdx.y with
// element strides sV (basis), sZ / sX (vectors), sY / sH (coefficients); partial buffers
// are per instance of nparts*KMAX (dots) or nparts (norms); active[b] == 0 skips instance b.
constexpr int KMAX = 128;  // max vectors in one orthogonalization
int krylov_partials_blocks(int n);
void launch_dot_partials(int is_float, const void *V, size_t ldv, int kcount, const void *z,
                         int n, double *P, hipStream_t s, int batch = 1, size_t sV = 0,
                         size_t sZ = 0, const int *active = nullptr);
// h[b*sH + k] = sum over chunks of the partials P (k < kcount)
void launch_reduce_parts(int kcount, const double *P, int n, double *h, hipStream_t s,
                         int batch = 1, size_t sH = 0, const int *active = nullptr);
// Pin: h is reduced here from the previous pass's chunk partials (h unused), workgroup 0
// storing it to hsave
void launch_update_dot(int is_float, const void *V, size_t ldv, int kcount, const double *h,
                       void *z, int n, double *Pout, double *normp, int want_norm, hipStream_t s,
                       int batch = 1, size_t sV = 0, size_t sZ = 0, size_t sH = 0,
                       const int *active = nullptr, const double *Pin = nullptr,
                       double *hsave = nullptr);
// hmap/hseq/seq: also publish hout (kcount + 1 values, single instance) to coherent pinned
// host memory, then seq to *hseq; xcl/pos (FP32, single instance): also xcl[pos[i]] = vout[i]
void launch_finalize(int is_float, int kcount, const double *h1, const double *h2,
                     const double *Pn, void *z, int n, void *vout, double *hout, hipStream_t s,
                     int batch = 1, size_t sZ = 0, size_t sV = 0, size_t sHin = 0,
                     size_t sH = 0, const int *active = nullptr, double *hmap = nullptr,
                     unsigned long long *hseq = nullptr, unsigned long long seq = 0,
                     float *xcl = nullptr, const int *pos = nullptr);
void launch_axpy_basis(int is_float, const void *V, size_t ldv, int kcount, const double *y,
                       void *x_acc, int x_is_float, int n, hipStream_t s, int batch = 1,
                       size_t sV = 0, size_t sY = 0, size_t sX = 0, const int *active = nullptr);
void launch_norm2_partials(int is_float, const void *x, int n, double *P, hipStream_t s,
                           int batch = 1, size_t sX = 0, const int *active = nullptr);
void launch_vec_ops(int op, int n, const void *a, const void *b, void *c, double alpha,
                    hipStream_t s);
void launch_permute(int n, const int *idx, const double *src, double *dst, int gather,
                    hipStream_t s);

// tangent of the residual: out[i] = sum_d dR_i/dx_d . dX_d (0 on rows flagged in skip);
// buffers: cstate nc*nq*2, trec nc*nq*8, part vjp_row_chunks(nc)*n, ds n, out n
void launch_residual_jvp(int q_order, int n, int nc, const double *X, const double *dX,
                         const int *row_cls, const double *phi, const double *psi,
                         const double *cellq, const int *cell_int, const int *cell_cls,
                         const int8_t *orient, const double *qpts, const double *qw,
                         const int *sing_ptr, const int *sing_cell, const int *sing_vtx, int ns,
                         const double *srule, const uint8_t *skip, double *cstate,
                         double *trec, double *part, double *ds, double *out, hipStream_t s);

}  // namespace wb
