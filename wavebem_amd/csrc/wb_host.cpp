// wb_host.cpp -- host-only C-ABI utilities (no device needed).
//
// They expose the library's own quadrature tables, double-node sets and constraint rules
// so that a caller (or a test without a GPU) can inspect exactly what the device path uses.
#include <algorithm>
#include <cstring>

#include "wb_internal.h"

extern "C" {

int wb_host_quadrature(int q_order, int s_order, double *reg_pts, double *reg_w,
                       double *sing_pts, double *sing_w)
{
    if (q_order < 1 || q_order > wb::MAX_NQ1 || s_order < 1 || s_order > wb::MAX_SING)
        return WB_E_ARG;
    wb::Quadrature Q;
    wb::build_quadrature(q_order, s_order, Q);
    if (reg_pts) std::memcpy(reg_pts, Q.reg_pts.data(), sizeof(double) * Q.reg_pts.size());
    if (reg_w) std::memcpy(reg_w, Q.reg_w.data(), sizeof(double) * Q.reg_w.size());
    if (sing_pts) std::memcpy(sing_pts, Q.sing_pts.data(), sizeof(double) * Q.sing_pts.size());
    if (sing_w) std::memcpy(sing_w, Q.sing_w.data(), sizeof(double) * Q.sing_w.size());
    return WB_OK;
}

int wb_host_mesh_plan(int n, int nc, const double *support, const int32_t *cell_dofs,
                      const int32_t *cell_material, int32_t *perm, int32_t *dn_ptr,
                      int32_t *dn_idx, int dn_cap, int32_t *n_tiles, int32_t *n_tile_cells,
                      int32_t *n_singular_pairs)
{
    if (n <= 0 || nc <= 0 || !support || !cell_dofs) return WB_E_ARG;
    wb::MeshPlan P;
    std::string e = wb::plan_mesh(n, nc, support, cell_dofs, cell_material, nullptr, nullptr,
                                  wb::TILE_COLS, 1, P);
    if (!e.empty()) return WB_E_ARG;
    if (perm) std::copy(P.perm.begin(), P.perm.end(), perm);
    if (dn_ptr) std::copy(P.dn_ptr.begin(), P.dn_ptr.end(), dn_ptr);
    if (dn_idx) {
        if ((int)P.dn_idx.size() > dn_cap) return WB_E_ARG;
        std::copy(P.dn_idx.begin(), P.dn_idx.end(), dn_idx);
    }
    if (n_tiles) *n_tiles = P.n_tasks + P.n_tiles;
    if (n_tile_cells) *n_tile_cells = (int32_t)(P.strip_cells + (long)P.tile_cell.size());
    if (n_singular_pairs) *n_singular_pairs = (int32_t)P.sing_cell.size();
    return WB_OK;
}

int wb_host_strip_plan(int n, int nc, const double *support, const int32_t *cell_dofs,
                       const int32_t *cell_material, int strips, int32_t *sizes, int32_t *tasks,
                       int32_t *gaps, int32_t *gap_cells, int32_t *cell_int_s, int32_t *perm)
{
    if (n <= 0 || nc <= 0 || !support || !cell_dofs || !sizes) return WB_E_ARG;
    wb::MeshPlan P;
    std::string e = wb::plan_mesh(n, nc, support, cell_dofs, cell_material, nullptr, nullptr,
                                  wb::TILE_COLS, strips ? 1 : 0, P);
    if (!e.empty()) return WB_E_ARG;
    sizes[0] = P.n_tasks;
    sizes[1] = (int32_t)(P.st_gap.size() / 4);
    sizes[2] = (int32_t)P.strip_cells;
    sizes[3] = P.strip_dofs;
    sizes[4] = P.n_tiles;
    if (tasks) std::copy(P.st_task.begin(), P.st_task.end(), tasks);
    if (gaps) std::copy(P.st_gap.begin(), P.st_gap.end(), gaps);
    if (gap_cells) std::copy(P.st_gcell.begin(), P.st_gcell.end(), gap_cells);
    if (cell_int_s) std::copy(P.cell_int_s.begin(), P.cell_int_s.end(), cell_int_s);
    if (perm) std::copy(P.perm.begin(), P.perm.end(), perm);
    return WB_OK;
}

int wb_host_cluster_plan(int n, int nc, const double *support, const int32_t *cell_dofs,
                         const int32_t *cell_material, int32_t *n_clusters, int32_t *cl_size,
                         int32_t *cl_idx, uint8_t *near, int32_t *perm)
{
    if (n <= 0 || nc <= 0 || !support || !cell_dofs || !n_clusters) return WB_E_ARG;
    wb::MeshPlan P;
    std::string e = wb::plan_mesh(n, nc, support, cell_dofs, cell_material, nullptr, nullptr,
                                  wb::TILE_COLS, 1, P);
    if (!e.empty()) return WB_E_ARG;
    std::vector<double> Xi(3 * (size_t)n);
    for (int d = 0; d < n; ++d)
        for (int k = 0; k < 3; ++k) Xi[3 * (size_t)P.perm[d] + k] = support[3 * (size_t)d + k];
    wb::HmatPlan H;
    wb::plan_hmat(P, Xi.data(), H);
    *n_clusters = H.ncl;
    if (cl_size) std::copy(H.cl_size.begin(), H.cl_size.end(), cl_size);
    if (cl_idx) std::copy(H.cl_idx.begin(), H.cl_idx.end(), cl_idx);
    if (near) std::copy(H.near.begin(), H.near.end(), near);
    if (perm) std::copy(P.perm.begin(), P.perm.end(), perm);
    return WB_OK;
}

int wb_host_hm_work_items(int ncl, const uint8_t *rank, const int32_t *cl_size,
                          int32_t *sizes, int32_t *item_ptr, int32_t *items)
{
    if (ncl <= 0 || !rank || !cl_size || !sizes) return WB_E_ARG;
    for (int J = 0; J < ncl; ++J)
        if (cl_size[J] < 0 || cl_size[J] > wb::HM_CL) return WB_E_ARG;
    for (long b = 0; b < (long)ncl * ncl; ++b)
        if (rank[b] != wb::HM_DENSE && rank[b] > wb::HM_RMAX) return WB_E_ARG;
    std::vector<int> ptr, it;
    wb::hm_plan_items(ncl, rank, cl_size, ptr, it);
    const int nitems = ptr.back();
    sizes[0] = wb::hm_chunks(ncl);
    sizes[1] = nitems;
    sizes[2] = wb::HM_DCOLS;
    sizes[3] = wb::HM_ITEM_SHIFT;
    sizes[4] = wb::HM_ITEM_LOWRANK;
    if (item_ptr) std::copy(ptr.begin(), ptr.end(), item_ptr);
    if (items) std::copy(it.begin(), it.begin() + nitems, items);
    return WB_OK;
}

int wb_host_compute_constraints(int n, const int32_t *dn_ptr, const int32_t *dn_idx,
                                const double *surface, const double *normals,
                                const double *grads, const double *tmp_rhs, int cap_lines,
                                int cap_nnz, int32_t *line_rows, int32_t *line_ptr,
                                int32_t *ent_col, double *ent_w, double *inhom)
{
    if (n <= 0 || !dn_ptr || !dn_idx || !surface || !tmp_rhs) return -WB_E_ARG;
    wb::ConstraintSet cs;
    std::string e =
        wb::build_constraints(n, dn_ptr, dn_idx, surface, normals, grads, tmp_rhs, cs);
    if (!e.empty()) return -WB_E_UNSUPPORTED;
    const int nl = (int)cs.rows.size();
    if (nl > cap_lines || (int)cs.cols.size() > cap_nnz) return -WB_E_ARG;
    std::copy(cs.rows.begin(), cs.rows.end(), line_rows);
    std::copy(cs.ptr.begin(), cs.ptr.end(), line_ptr);
    std::copy(cs.cols.begin(), cs.cols.end(), ent_col);
    std::copy(cs.w.begin(), cs.w.end(), ent_w);
    std::copy(cs.inhom.begin(), cs.inhom.end(), inhom);
    return nl;
}

}  // extern "C"
