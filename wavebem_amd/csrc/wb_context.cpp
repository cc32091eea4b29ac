// wb_context.cpp -- C-ABI entry points, device state and the GMRES driver.
//
// Stands in for the method bodies of BEMProblem<3> (source/bem_problem.cc); each entry
// point's reference counterpart is named in include/wavebem_bem.h.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <functional>
#include <vector>

#include "wb_internal.h"

using namespace wb;

namespace {

constexpr int HMAP_STRIDE = 136;  // doubles per GMRES column slot in pinned memory
static_assert(KMAX + 1 < HMAP_STRIDE, "a column and its sequence word fit one slot");

struct WbError : std::runtime_error {
    wb_status code;
    WbError(wb_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw WbError(e_ == hipErrorOutOfMemory ? WB_E_OOM : WB_E_HIP,                   \
                          std::string(#expr) + ": " + hipGetErrorString(e_));                \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t b)
    {
        if (b <= bytes && p) return;
        release();
        if (b == 0) return;
        HIPCHK(hipMalloc(&p, b));
        bytes = b;
    }
    template <typename T> T *as() const { return (T *)p; }
};

double now_ms()
{
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

struct wb_ctx {
    wb_config cfg{};
    hipStream_t stream = nullptr;
    // second stream for preconditioner builds that overlap the compressed operator's build
    // (default priority), and the events that order it against the main stream
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_main = nullptr, ev_side = nullptr;
    std::string err;
    Quadrature quad;
    MeshPlan plan;
    bool have_mesh = false, have_masks = false, assembled = false;
    int n = 0;
    size_t ld = 0;
    std::vector<double> surface, other;  // caller order
    std::vector<int> cell_dofs_h, cell_mat_h;  // caller numbering; materials may be empty
    bool pure_neumann = false;
    // constraints (caller numbering) and whether they are rebuilt from tmp_rhs
    ConstraintSet cons;
    bool cons_auto = false;
    std::vector<double> cons_normals, cons_grads;
    // rebuilt constraints take normals / surface gradients from the device projections
    // when the caller passed none
    bool cons_project = false;
    bool have_cons = false;
    // versions for the preconditioner cache
    long assembly_ver = 0, mask_ver = 0, cons_struct_ver = 0;
    long pc_assembly_ver = -1, pc_mask_ver = -1, pc_cons_ver = -1;
    int pc_bs = 0;
    // coarse-space correction of the block-Jacobi preconditioner (wb_coarse.hip)
    bool cs_on = false;
    bool cs_t_on = false;  // the transposed-system data below match the current build
    int cs_agg = 0, cs_nc = 0;
    DevBuf cs_AP, cs_Ac, cs_Aci, cs_Aci32, cs_WT, cs_WT32, cs_status, cs_C, cs_T;
    DevBuf cs_PTA, cs_WTT;  // transposed system (adjoint solve), FP64
    DevBuf cs_S;            // aggregate row sums of the FP16 copy (nc x vlen, FP64)

    // device: mesh
    DevBuf tile_start;
    DevBuf X, cell_int, orient, cell_cls, row_cls, tile_ptr, tile_cell, tile_cols, sing_ptr,
        sing_cell, sing_vtx, qpts, qw, srule, cellq, perm_d, iperm_d;
    // strip schedule of the regular-cell assembly and its cell records in gap order
    // (canonical vertex order of the structured patches)
    DevBuf st_task, st_gap, st_gcell, cell_int_s, orient_s, cellq_s;
    // device: operator
    DevBuf G, H, Glo, Hlo, hsum_part, hsum_sing, alpha;
    DevBuf surf_d, gmask, diag_op, diag_rhs, segs;
    int n_segs = 0;
    DevBuf con_rows, con_ptr, con_cols, con_w, con_inhom, constrained, free_rows;
    DevBuf con_line;  // per row: its constraint line, -1 for free rows
    DevBuf surf_c;    // surface mask in caller order (uint8)
    DevBuf cs_Cpart;  // coarse restriction partials of the compressed product (64-row blocks)
    // transposed constraint lines, gathered by target column: y[tgt] += coef * x[src]
    DevBuf ct_tgt, ct_ptr, ct_src, ct_coef;
    int n_ct = 0;
    bool op_transpose = false;  // apply_constrained / apply_precond act with A^T
    DevBuf gt_part;             // chunk partials of the transposed product
    // backward of the assembly
    DevBuf vjp_w, vjp_phi, vjp_psi, vjp_rows, vjp_cstate, vjp_p1, vjp_p2, vjp_xg, vjp_vbar,
        vjp_xs, vjp_spart;
    // L2 projections: node -> (cell, local vertex) incidence, element blocks, CG vectors
    DevBuf inc_ptr, inc_item, pj_emass, pj_erhs, pj_u, pj_psi, pj_B, pj_X, pj_R, pj_Z, pj_P,
        pj_Q, pj_dinv, pj_part, pj_part2, pj_state;
    int proj_iterations = 0;
    int n_lines = 0, n_free = 0;
    DevBuf binv, binv32;
    // row-scaled FP16 copy of the FP32 vmult operator for the inner solves
    DevBuf A16, rs16;
    long h16_assembly_ver = -1, h16_mask_ver = -1;
    int h16_agg = 0;        // aggregate size the FP16 copy was built for (0: none)
    bool h16_sums = false;  // cs_S holds its aggregate sums
    // compressed copy of the FP32 vmult operator (ir_inner_format 2, wb_hmat.hip): the
    // cluster plan of the mesh, then per assembly and masks the block ranks, scales and
    // factors
    HmatPlan hmp;
    DevBuf hm_idx, hm_size, hm_seg, hm_near, hm_rank, hm_scale, hm_Y, hm_Mx, hm_Q, hm_B, hm_D,
        hm_doff, hm_T, hm_part, hm_iptr, hm_items, hm_pos;
    // the FP32 vector whose cluster-order copy hm_T holds (written by the GMRES step that
    // produced it), or nullptr
    const void *hm_xc_for = nullptr;
    long hm_assembly_ver = -1, hm_mask_ver = -1;
    int hm_agg = 0;         // aggregate size of the coarse sums written by the build (0: none)
    int float_mode = 1;  // operator mode of the FP32-vector solves: 1 = FP32 hi, 3 = FP16,
                         // 4 = compressed
    // device: solver work
    DevBuf V, z, w, P1, P2, Pn, hbuf, hred, vx, vb, vr, vt, vs;   // double / raw
    DevBuf V32, z32, w32, r32, d32;
    double *hpin = nullptr;  // pinned host copy of hbuf
    // coherent pinned host memory the GMRES columns are published to: two slots of
    // HMAP_STRIDE doubles (up to KMAX + 1 values, the sequence number in the last word),
    // publication seq using slot seq & 1; device-side address of the same memory
    double *hmap = nullptr, *hmap_dev = nullptr;
    unsigned long long hmap_seq = 0;
    size_t vlen = 0;         // padded vector length

    // kernel timing
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    unsigned long long kcalls[WB_K_COUNT] = {};  // launches seen per kind while profiling
    struct Pending {
        int kind;
        size_t ev;
        long long rows;
    };
    std::vector<Pending> pending;
    wb_kernel_stats kstats{};

    // several right-hand sides at once
    struct Multi {
        int k = 0;
        DevBuf V, V32, Z, W, Z32, W32, P1, P2, Pn, H, hred, X, B, R, T, D32, R32, act, inh, part,
            xs, hxc, hpart;
        double *hpin = nullptr;  // pinned k x 132 coefficients
        int *apin = nullptr;     // pinned k active flags
        std::vector<DevBuf *> all()
        {
            return {&V, &V32, &Z, &W, &Z32, &W32, &P1, &P2, &Pn, &H, &hred,
                    &X, &B, &R, &T, &D32, &R32, &act, &inh, &part, &xs, &hxc, &hpart};
        }
    } mw;

    // every device buffer of the context
    std::vector<DevBuf *> all_bufs()
    {
        std::vector<DevBuf *> v = {
            &X, &cell_int, &orient, &cell_cls, &row_cls, &tile_ptr, &tile_cell, &tile_cols,
            &tile_start, &sing_ptr, &sing_cell, &sing_vtx, &qpts, &qw, &srule, &cellq, &perm_d,
            &iperm_d, &st_task, &st_gap, &st_gcell, &cell_int_s, &orient_s, &cellq_s, &G, &H,
            &Glo, &Hlo, &hsum_part, &hsum_sing, &alpha, &surf_d, &gmask, &segs, &diag_op,
            &diag_rhs, &con_rows, &con_ptr, &con_cols, &con_w, &con_inhom, &constrained,
            &con_line, &cs_Cpart, &surf_c,
            &free_rows, &binv, &binv32, &cs_AP, &cs_Ac, &cs_Aci, &cs_Aci32, &cs_WT,
            &cs_WT32, &cs_status, &cs_C, &cs_T, &cs_PTA, &cs_WTT, &cs_S, &V, &z, &w, &P1, &P2,
            &Pn, &hbuf, &hred, &vx, &vb, &vr, &vt, &vs, &V32, &z32, &w32, &r32, &d32, &ct_tgt,
            &ct_ptr, &ct_src, &ct_coef, &gt_part, &vjp_w, &vjp_phi, &vjp_psi, &vjp_rows,
            &vjp_cstate, &vjp_p1, &vjp_p2, &vjp_xg, &vjp_vbar, &vjp_xs, &vjp_spart, &inc_ptr,
            &inc_item, &pj_emass, &pj_erhs, &pj_u, &pj_psi, &pj_B, &pj_X, &pj_R, &pj_Z, &pj_P,
            &pj_Q, &pj_dinv, &pj_part, &pj_part2, &pj_state, &A16, &rs16, &hm_idx, &hm_size, &hm_seg,
            &hm_near, &hm_rank, &hm_scale, &hm_Y, &hm_Mx, &hm_Q, &hm_B, &hm_D, &hm_doff, &hm_T,
            &hm_part, &hm_iptr, &hm_items, &hm_pos};
        for (DevBuf *b : mw.all()) v.push_back(b);
        return v;
    }
};

// Brackets one launch with a pair of events when profiling is on.  The products launched
// once per GMRES step (FP64, FP32 and FP16 matrices, compressed operator) are timed on one
// launch in STEP_SAMPLE, in both precisions alike: each timing event in the stream costs
// about 5.6 us of idle GPU, against 70 us for the compressed product.  calls counts them all.
constexpr unsigned STEP_SAMPLE = 8;
static bool step_product(int kind)
{
    return kind == WB_K_GEMV64 || kind == WB_K_GEMV32 || kind == WB_K_GEMV16 ||
           kind == WB_K_GEMVC;
}
struct KernelScope {
    wb_ctx *c;
    int kind;
    long long rows;
    size_t ev = 0;
    bool on = false;
    KernelScope(wb_ctx *c_, int kind_, long long rows_ = 0) : c(c_), kind(kind_), rows(rows_)
    {
        if (!c->profiling) return;
        c->kstats.calls[kind] += 1;
        if (step_product(kind) && c->kcalls[kind]++ % STEP_SAMPLE != 0) return;
        on = true;
        while (c->ev_pool.size() < c->ev_used + 2) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            c->ev_pool.push_back(e);
        }
        ev = c->ev_used;
        c->ev_used += 2;
        HIPCHK(hipEventRecord(c->ev_pool[ev], c->stream));
    }
    ~KernelScope()
    {
        if (!on) return;
        if (hipEventRecord(c->ev_pool[ev + 1], c->stream) == hipSuccess)
            c->pending.push_back({kind, ev, rows});
    }
};

static void fold_kernel_stats(wb_ctx *c)
{
    if (c->pending.empty()) return;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (const auto &p : c->pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev_pool[p.ev], c->ev_pool[p.ev + 1]));
        c->kstats.launches[p.kind] += 1;
        c->kstats.total_ms[p.kind] += ms;
        c->kstats.rows[p.kind] += p.rows;
    }
    c->pending.clear();
    c->ev_used = 0;
}

static void set_err(wb_ctx *c, const std::string &m)
{
    if (c) c->err = m;
}

template <typename F>
static wb_status guarded(wb_ctx *c, F &&f)
{
    if (!c) return WB_E_ARG;
    try {
        c->err.clear();
        f();
        return WB_OK;
    } catch (const WbError &e) {
        set_err(c, e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_err(c, "host allocation failed");
        return WB_E_OOM;
    } catch (const std::exception &e) {
        set_err(c, e.what());
        return WB_E_HIP;
    }
}

static void require(bool cond, wb_status code, const char *msg)
{
    if (!cond) throw WbError(code, msg);
}

template <typename T>
static void upload(DevBuf &b, const T *src, size_t count, hipStream_t s)
{
    b.ensure(sizeof(T) * std::max<size_t>(count, 1));
    if (count) HIPCHK(hipMemcpyAsync(b.p, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
}

static size_t elem_bytes(const wb_ctx *c) { return c->cfg.precision == WB_PREC_FP64 ? 8 : 4; }

// Row-sum partials of H written by the regular-cell kernels: one per strip task, then one
// per group of TILES_PER_WG column tiles.
static int hsum_parts(const wb_ctx *c)
{
    return c->plan.n_tasks + (c->plan.n_tiles + TILES_PER_WG - 1) / TILES_PER_WG;
}

// Quadrature-point records at the current support points: cellq for every cell in its
// own vertex order (singular correction, adjoint, tangent) and cellq_s for every gap
// position of the strip schedule, in the canonical vertex order.
static void cell_records(wb_ctx *c)
{
    const int nc = c->plan.nc, nq = c->quad.q_order * c->quad.q_order;
    launch_cell_geometry(c->X.as<double>(), c->cell_int.as<int>(), c->orient.as<int8_t>(), nc,
                         nq, c->qpts.as<double>(), c->qw.as<double>(), c->cellq.as<double>(),
                         c->stream);
    HIPCHK(hipGetLastError());
    if (c->plan.n_tasks > 0) {
        launch_cell_geometry(c->X.as<double>(), c->cell_int_s.as<int>(),
                             c->orient_s.as<int8_t>(), (int)c->plan.st_gcell.size(), nq,
                             c->qpts.as<double>(), c->qw.as<double>(), c->cellq_s.as<double>(),
                             c->stream, c->st_gcell.as<int>());
        HIPCHK(hipGetLastError());
    }
}
static void alloc_matrices(wb_ctx *c);
static void ensure_h16(wb_ctx *c);
static void ensure_hm(wb_ctx *c, const std::function<void()> &after_ranks = {});

// coarse aggregates: a multiple of 8 dofs, at least 64, at most precond_coarse of them
static void coarse_shape(const wb_ctx *c, int &agg, int &nc)
{
    const int maxc = c->cfg.precond_coarse;
    agg = std::max(64, 8 * ((c->n + 8 * maxc - 1) / (8 * maxc)));
    nc = (c->n + agg - 1) / agg;
}

// ------------------------------------------------------------------ vector plumbing
static void to_internal(const wb_ctx *c, const double *src, std::vector<double> &dst)
{
    dst.assign(c->vlen, 0.0);
    for (int d = 0; d < c->n; ++d) dst[c->plan.perm[d]] = src[d];
}

static void to_caller(const wb_ctx *c, const std::vector<double> &src, double *dst)
{
    for (int d = 0; d < c->n; ++d) dst[d] = src[c->plan.perm[d]];
}

static void upload_vec(wb_ctx *c, DevBuf &b, const double *caller)
{
    std::vector<double> tmp;
    to_internal(c, caller, tmp);
    b.ensure(sizeof(double) * c->vlen);
    HIPCHK(hipMemcpyAsync(b.p, tmp.data(), sizeof(double) * c->vlen, hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
}

static void download_vec(wb_ctx *c, const DevBuf &b, double *caller)
{
    std::vector<double> tmp(c->vlen);
    HIPCHK(hipMemcpyAsync(tmp.data(), b.p, sizeof(double) * c->vlen, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    to_caller(c, tmp, caller);
}

static void ensure_work(wb_ctx *c)
{
    const size_t vb = sizeof(double) * c->vlen;
    for (DevBuf *b : {&c->z, &c->w, &c->vx, &c->vb, &c->vr, &c->vt, &c->vs}) {
        if (b->bytes < vb) {
            b->ensure(vb);
            HIPCHK(hipMemsetAsync(b->p, 0, vb, c->stream));
        }
    }
    const int parts = krylov_partials_blocks(c->n);
    c->P1.ensure(sizeof(double) * parts * 128);
    c->P2.ensure(sizeof(double) * parts * 128);
    c->Pn.ensure(sizeof(double) * parts);
    c->hbuf.ensure(sizeof(double) * 132);
    c->hred.ensure(sizeof(double) * 2 * KMAX);
    if (!c->hpin) HIPCHK(hipHostMalloc((void **)&c->hpin, sizeof(double) * 132, 0));
    if (!c->hmap) {
        HIPCHK(hipHostMalloc((void **)&c->hmap, sizeof(double) * 2 * HMAP_STRIDE,
                             hipHostMallocMapped | hipHostMallocCoherent));
        HIPCHK(hipHostGetDevicePointer((void **)&c->hmap_dev, c->hmap, 0));
        for (int slot = 0; slot < 2; ++slot)
            *(volatile unsigned long long *)(c->hmap + slot * HMAP_STRIDE + HMAP_STRIDE - 1) = 0;
        c->hmap_seq = 0;
    }
}

// Launch of k_finalize that also publishes the column as publication seq
static void finalize_publish(wb_ctx *c, int is_float, int kcount, const double *h1,
                             const double *h2, void *z, int n, void *vout,
                             unsigned long long seq)
{
    double *slot = c->hmap_dev + (seq & 1) * HMAP_STRIDE;
    // with the compressed inner operator, the new basis vector also goes to the product's
    // cluster-order copy, so its product needs no gather
    const bool xc = is_float && c->float_mode == 4 && c->hm_T.bytes >= sizeof(float) *
                                                                         (size_t)c->hmp.ncl * HM_CL;
    launch_finalize(is_float, kcount, h1, h2, c->Pn.as<double>(), z, n, vout,
                    c->hbuf.as<double>(), c->stream, 1, 0, 0, 0, 0, nullptr, slot,
                    (unsigned long long *)(slot + HMAP_STRIDE - 1), seq,
                    xc ? c->hm_T.as<float>() : nullptr, xc ? c->hm_pos.as<int>() : nullptr);
    c->hm_xc_for = xc ? vout : nullptr;
}

// The device's publication of a GMRES column (k_finalize with hmap): waits until the
// sequence number reaches seq.  The stream is queried now and then, so an error or an
// idle stream that never published ends the wait with an exception instead of a hang.
static const double *wait_hmap(wb_ctx *c, unsigned long long seq)
{
    const double *slot = c->hmap + (seq & 1) * HMAP_STRIDE;
    const unsigned long long *p = (const unsigned long long *)(slot + HMAP_STRIDE - 1);
    for (unsigned spins = 1;; ++spins) {
        if (__atomic_load_n(p, __ATOMIC_ACQUIRE) >= seq) return slot;
        if ((spins & 1023u) == 0) {
            const hipError_t e = hipStreamQuery(c->stream);
            if (e != hipSuccess && e != hipErrorNotReady) HIPCHK(e);
            if (e == hipSuccess) {
                if (__atomic_load_n(p, __ATOMIC_ACQUIRE) >= seq) return slot;
                throw WbError(WB_E_HIP, "GMRES column was not published");
            }
        }
    }
}

static double device_norm(wb_ctx *c, const void *x, int is_float)
{
    const int parts = krylov_partials_blocks(c->n);
    launch_norm2_partials(is_float, x, c->n, c->Pn.as<double>(), c->stream);
    std::vector<double> p(parts);
    HIPCHK(hipMemcpyAsync(p.data(), c->Pn.p, sizeof(double) * parts, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (double v : p) s += v;
    return std::sqrt(s);
}

// ------------------------------------------------------------------ operator pieces
// which: 0 = BEMProblem::vmult selection, 1 = compute_rhs selection
// The compressed product y = A~ x + alpha .* x (mode 4); with lines, constrained rows get
// their constraint line's value in the same launch (apply_constrained); with cparts, the
// last kernel also leaves the coarse restriction partials of y in cs_Cpart.
static void hm_product(wb_ctx *c, const void *x, void *y, bool lines, bool cparts)
{
    KernelScope ks(c, WB_K_GEMVC, c->n);
    const bool ready = x == c->hm_xc_for;
    c->hm_xc_for = nullptr;
    if (cparts) c->cs_Cpart.ensure(sizeof(double) * 2 * (size_t)((c->n + 63) / 64));
    launch_hm_gemv(c->hm_idx.as<int>(), c->hm_size.as<int>(), c->hmp.ncl,
                   c->hm_rank.as<uint8_t>(), c->hm_scale.as<float>(), c->hm_Q.p, c->hm_B.p,
                   c->hm_D.p, c->hm_doff.as<int>(), c->hm_iptr.as<int>(), c->hm_items.as<int>(),
                   c->n, (const float *)x, c->diag_op.as<double>(), (const float *)x,
                   c->hm_T.as<float>(), c->hm_part.as<float>(), c->vlen, (float *)y, c->stream,
                   ready, lines && c->n_lines > 0 ? c->con_line.as<int>() : nullptr,
                   c->con_ptr.as<int>(), c->con_cols.as<int>(), c->con_w.as<double>(),
                   cparts ? c->cs_agg : 64, cparts ? c->cs_Cpart.as<double>() : nullptr);
}

static void gemv(wb_ctx *c, int mode, int which, const void *x, const void *xd, void *y,
                 bool free_rows_only)
{
    const double *diag = which == 0 ? c->diag_op.as<double>() : c->diag_rhs.as<double>();
    const double s_sel = which == 0 ? 1.0 : -1.0, s_uns = which == 0 ? -1.0 : 1.0;
    const int *rows = free_rows_only ? c->free_rows.as<int>() : nullptr;
    const int nr = free_rows_only ? c->n_free : c->n;
    if (mode == 4) {
        require(which == 0, WB_E_STATE, "the compressed operator holds the vmult selection only");
        require(xd == x, WB_E_STATE, "the compressed product takes its diagonal term from x");
        hm_product(c, x, y, false, false);
        return;
    }
    if (mode == 3) {
        require(which == 0, WB_E_STATE, "the FP16 operator holds the vmult selection only");
        KernelScope ks(c, WB_K_GEMV16, nr);
        launch_gemv_h16(c->A16.p, c->rs16.as<float>(), c->ld, c->n, (const float *)x, diag,
                        (const float *)xd, rows, nr, (float *)y, c->stream);
        return;
    }
    KernelScope ks(c, mode == 0 ? WB_K_GEMV64 : (mode == 1 ? WB_K_GEMV32 : WB_K_GEMVHL), nr);
    launch_gemv_sel(mode, c->G.p, c->H.p, c->Glo.p, c->Hlo.p, c->ld, c->segs.as<int>(),
                    c->n_segs, c->gmask.as<uint8_t>(), which, s_sel, s_uns, x, diag, xd, rows, nr,
                    y, c->stream);
}

// y = ConstrainedOperator(A) x in FP64 (mode 0 or 2) or FP32 vectors (mode 1: FP32
// operator, mode 3: FP16 operator, mode 4: compressed operator)
// whether the block-Jacobi apply can take C = P^T x from the compressed product's partials
static bool coarse_parts_usable(const wb_ctx *c)
{
    return c->cfg.preconditioner == WB_PRECOND_BLOCK_JACOBI && c->cs_on && !c->op_transpose &&
           bj_coarse_fits(c->pc_bs, c->cs_agg, c->cs_nc);
}

// y = A_c x.  With coarse_parts and the compressed operator, the product also leaves the
// coarse restriction partials of y in cs_Cpart; returns whether it did (the caller then
// hands them to the preconditioner of the same y, apply_precond(..., true)).
static bool apply_constrained(wb_ctx *c, int mode, const void *x, void *y,
                              bool coarse_parts = false)
{
    const int is_float = mode == 1 || mode == 3 || mode == 4;
    if (c->op_transpose) {
        // A_c^T x: transposed product over the unconstrained rows (the diagonal alpha term
        // of those rows included), then the transposed constraint lines
        // (mode 0: FP64 matrices; FP32-IR: mode 1 the hi part on FP32 vectors, mode 2 hi + lo)
        require(mode <= 2, WB_E_UNSUPPORTED,
                "the transposed operator has no FP16 or compressed form");
        const int chunks = gemvT_chunks(c->n);
        c->gt_part.ensure(sizeof(double) * c->ld * chunks);
        KernelScope ks(c, mode == 0 ? WB_K_GEMV64 : (mode == 1 ? WB_K_GEMV32 : WB_K_GEMVHL),
                       (long long)c->n_free);
        launch_gemvT(mode, c->G.p, c->H.p, c->Glo.p, c->Hlo.p, c->ld, c->n,
                     c->gmask.as<uint8_t>(), 0, 1.0, -1.0, x, c->diag_op.as<double>(),
                     c->constrained.as<uint8_t>(), c->gt_part.as<double>(), y, c->stream);
        launch_constraint_T(is_float, c->n_ct, c->ct_tgt.as<int>(), c->ct_ptr.as<int>(),
                            c->ct_src.as<int>(), c->ct_coef.as<double>(), x, y, c->stream);
        return false;
    }
    if (mode == 4 && !c->pure_neumann) {
        // the constraint lines (and the partials) in the product's last kernel
        const bool parts = coarse_parts && coarse_parts_usable(c);
        hm_product(c, x, y, true, parts);
        return parts;
    }
    if (c->pure_neumann) {
        gemv(c, mode, 0, x, x, y, false);
        const double nrm = device_norm(c, y, is_float);
        // dst.add(-dst.l2_norm()): add -||dst|| to every entry
        std::vector<char> tmp;
        if (is_float) {
            std::vector<float> v(c->n);
            HIPCHK(hipMemcpy(v.data(), y, sizeof(float) * c->n, hipMemcpyDeviceToHost));
            for (auto &e : v) e = (float)((double)e - nrm);
            HIPCHK(hipMemcpy(y, v.data(), sizeof(float) * c->n, hipMemcpyHostToDevice));
        } else {
            std::vector<double> v(c->n);
            HIPCHK(hipMemcpy(v.data(), y, sizeof(double) * c->n, hipMemcpyDeviceToHost));
            for (auto &e : v) e -= nrm;
            HIPCHK(hipMemcpy(y, v.data(), sizeof(double) * c->n, hipMemcpyHostToDevice));
        }
    } else {
        gemv(c, mode, 0, x, x, y, true);
    }
    launch_constraint_rows(is_float, c->n_lines, c->con_rows.as<int>(), c->con_ptr.as<int>(),
                           c->con_cols.as<int>(), c->con_w.as<double>(),
                           c->con_inhom.as<double>(), 0, x, y, c->stream);
    return false;
}

// Block Jacobi, with the coarse-space correction (forward operator only):
// y = P z + B^-1 (x - AP z), z = Ac^-1 P^T x.  x and y may alias.  cparts: cs_Cpart holds
// the restriction partials of x (apply_constrained with coarse_parts returned true).
static void block_precond(wb_ctx *c, int is_float, int k, const void *x, void *y,
                          bool cparts = false)
{
    const void *binv = is_float ? c->binv32.p : c->binv.p;
    if (c->cs_on && c->op_transpose && c->cs_t_on && !is_float) {
        // the transposed system: y = P z + B^-T (x - A^T P z), z = Ac^-T P^T x
        launch_coarse_pre(0, c->n, c->cs_agg, c->cs_nc, x, c->vlen, c->cs_WTT.p, c->vlen,
                          c->cs_C.as<double>(), c->cs_T.p, k, c->stream);
        launch_block_jacobi_apply(0, c->n, c->pc_bs, binv, c->cs_T.p, y, k, c->vlen, c->stream,
                                  1, c->cs_C.as<double>(), c->cs_Aci.p, c->cs_nc, c->cs_agg);
    } else if (c->cs_on && !c->op_transpose && k == 1 &&
               bj_coarse_fits(c->pc_bs, c->cs_agg, c->cs_nc)) {
        if (!cparts)
            launch_coarse_restrict(is_float, c->n, c->cs_agg, c->cs_nc, x, c->vlen,
                                   c->cs_C.as<double>(), 1, c->stream);
        launch_bj_coarse(is_float, c->n, c->pc_bs, binv, x, y,
                         is_float ? c->cs_WT32.p : c->cs_WT.p, c->vlen, c->cs_C.as<double>(),
                         is_float ? c->cs_Aci32.p : c->cs_Aci.p, c->cs_nc, c->cs_agg,
                         c->stream, cparts ? c->cs_Cpart.as<double>() : nullptr);
    } else if (c->cs_on && !c->op_transpose) {
        launch_coarse_pre(is_float, c->n, c->cs_agg, c->cs_nc, x, c->vlen,
                          is_float ? c->cs_WT32.p : c->cs_WT.p, c->vlen, c->cs_C.as<double>(),
                          c->cs_T.p, k, c->stream);
        launch_block_jacobi_apply(is_float, c->n, c->pc_bs, binv, c->cs_T.p, y, k, c->vlen,
                                  c->stream, 0, c->cs_C.as<double>(),
                                  is_float ? c->cs_Aci32.p : c->cs_Aci.p, c->cs_nc, c->cs_agg);
    } else {
        launch_block_jacobi_apply(is_float, c->n, c->pc_bs, binv, x, y, k, c->vlen, c->stream,
                                  c->op_transpose ? 1 : 0);
    }
}

static void apply_precond(wb_ctx *c, int is_float, const void *x, void *y, bool cparts = false)
{
    if (c->cfg.preconditioner == WB_PRECOND_BLOCK_JACOBI) {
        block_precond(c, is_float, 1, x, y, cparts);
    } else {
        HIPCHK(hipMemcpyAsync(y, x, (is_float ? 4 : 8) * (size_t)c->n, hipMemcpyDeviceToDevice,
                              c->stream));
    }
}

static void refresh_diag(wb_ctx *c)
{
    // diag_op = alpha .* other, diag_rhs = -alpha .* surface (internal order; other is
    // 1 - surface, surf_d the surface mask in internal order)
    c->diag_op.ensure(sizeof(double) * c->vlen);
    c->diag_rhs.ensure(sizeof(double) * c->vlen);
    launch_diag_from_mask(c->n, c->alpha.as<double>(), c->surf_d.as<double>(),
                          c->diag_op.as<double>(), c->diag_rhs.as<double>(), c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
}

static void install_constraints(wb_ctx *c, const ConstraintSet &cs)
{
    const int n = c->n;
    const int nl = (int)cs.rows.size();
    if (c->have_cons && c->cons.rows == cs.rows && c->cons.ptr == cs.ptr &&
        c->cons.cols == cs.cols && c->cons.w == cs.w) {
        // the installed lines (already checked) with, at most, new inhomogeneities: the
        // device copies of everything else stay
        if (c->cons.inhom != cs.inhom) {
            c->cons.inhom = cs.inhom;
            upload(c->con_inhom, c->cons.inhom.data(), c->cons.inhom.size(), c->stream);
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        return;
    }
    std::vector<int> rows(nl), ptr(cs.ptr), cols(cs.cols.size());
    std::vector<uint8_t> flag(n, 0);
    for (int l = 0; l < nl; ++l) {
        require(cs.rows[l] >= 0 && cs.rows[l] < n, WB_E_ARG, "constraint row out of range");
        rows[l] = c->plan.perm[cs.rows[l]];
        require(!flag[rows[l]], WB_E_ARG, "duplicate constraint line");
        flag[rows[l]] = 1;
    }
    for (size_t p = 0; p < cs.cols.size(); ++p) {
        require(cs.cols[p] >= 0 && cs.cols[p] < n, WB_E_ARG, "constraint column out of range");
        cols[p] = c->plan.perm[cs.cols[p]];
    }
    bool same_struct = c->have_cons && (int)c->cons.rows.size() == nl && c->cons.rows == cs.rows &&
                       c->cons.ptr == cs.ptr && c->cons.cols == cs.cols && c->cons.w == cs.w;
    c->cons = cs;
    c->have_cons = true;
    c->n_lines = nl;
    upload(c->con_rows, rows.data(), rows.size(), c->stream);
    upload(c->con_ptr, ptr.data(), ptr.size(), c->stream);
    upload(c->con_cols, cols.data(), cols.size(), c->stream);
    upload(c->con_w, cs.w.data(), cs.w.size(), c->stream);
    upload(c->con_inhom, cs.inhom.data(), cs.inhom.size(), c->stream);
    upload(c->constrained, flag.data(), flag.size(), c->stream);
    {
        std::vector<int> line(std::max(n, 1), -1);
        for (int l = 0; l < nl; ++l) line[rows[l]] = l;
        upload(c->con_line, line.data(), line.size(), c->stream);
    }
    std::vector<int> fr;
    fr.reserve(n);
    for (int i = 0; i < n; ++i)
        if (!flag[i]) fr.push_back(i);
    c->n_free = (int)fr.size();
    upload(c->free_rows, fr.data(), fr.size(), c->stream);
    // A_c^T adds, for line l with row r: +x[r] to y[r] and -w_p x[r] to y[col_p]
    {
        std::vector<std::vector<std::pair<int, double>>> by(n);
        for (int l = 0; l < nl; ++l) {
            by[rows[l]].push_back({rows[l], 1.0});
            for (int p = ptr[l]; p < ptr[l + 1]; ++p) by[cols[p]].push_back({rows[l], -cs.w[p]});
        }
        std::vector<int> tgt, tptr{0}, src;
        std::vector<double> coef;
        for (int j = 0; j < n; ++j) {
            if (by[j].empty()) continue;
            tgt.push_back(j);
            for (auto &e : by[j]) {
                src.push_back(e.first);
                coef.push_back(e.second);
            }
            tptr.push_back((int)src.size());
        }
        c->n_ct = (int)tgt.size();
        upload(c->ct_tgt, tgt.data(), tgt.size(), c->stream);
        upload(c->ct_ptr, tptr.data(), tptr.size(), c->stream);
        upload(c->ct_src, src.data(), src.size(), c->stream);
        upload(c->ct_coef, coef.data(), coef.size(), c->stream);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!same_struct) c->cons_struct_ver++;
}

static void ensure_precond(wb_ctx *c)
{
    if (c->cfg.preconditioner != WB_PRECOND_BLOCK_JACOBI) return;
    if (c->pc_assembly_ver == c->assembly_ver && c->pc_mask_ver == c->mask_ver &&
        c->pc_cons_ver == c->cons_struct_ver)
        return;
    const int bs = c->pc_bs;
    const int nb = (c->n + bs - 1) / bs;
    c->binv.ensure(sizeof(double) * (size_t)nb * bs * bs);
    c->binv32.ensure(sizeof(float) * (size_t)nb * bs * bs);
    // With the compressed inner operator, the blocks and the coarse space are built on the
    // second stream, beside the operator's sketch and fill on the main stream (they read
    // the matrices, the masks and the sketch's aggregate sums, and write only their own
    // buffers); the main stream then waits for them.
    const bool side = c->cfg.precision == WB_PREC_FP32_IR && c->cfg.ir_inner_format == 2 &&
                      c->cfg.precond_coarse > 0;
    hipStream_t bs_stream = c->stream;
    if (side) {
        HIPCHK(hipEventRecord(c->ev_main, c->stream));
        HIPCHK(hipStreamWaitEvent(c->stream2, c->ev_main, 0));
        bs_stream = c->stream2;
    }
    launch_block_jacobi_build(c->n, bs, c->G.p, c->H.p, c->Glo.p, c->Hlo.p, c->cfg.precision,
                              c->ld, c->alpha.as<double>(), c->surf_d.as<double>(),
                              c->constrained.as<uint8_t>(), c->binv.as<double>(),
                              c->binv32.as<float>(), bs_stream);
    HIPCHK(hipGetLastError());
    c->cs_on = false;
    c->cs_t_on = false;
    if (c->cfg.precond_coarse > 0) {
        int agg = 0, nc = 0;
        coarse_shape(c, agg, nc);
        c->cs_agg = agg;
        c->cs_nc = nc;
        c->cs_AP.ensure(sizeof(double) * (size_t)nc * c->vlen);
        c->cs_Ac.ensure(sizeof(double) * (size_t)nc * nc);
        c->cs_Aci.ensure(sizeof(double) * (size_t)nc * nc);
        c->cs_Aci32.ensure(sizeof(float) * (size_t)nc * nc);
        c->cs_status.ensure(sizeof(int));
        c->cs_WT.ensure(sizeof(double) * (size_t)nc * c->vlen);
        c->cs_WT32.ensure(sizeof(float) * (size_t)nc * c->vlen);
        c->cs_C.ensure(sizeof(double) * CS_MAX * WB_MULTI_MAX);
        c->cs_T.ensure(sizeof(double) * c->vlen * WB_MULTI_MAX);
        // FP32-IR with the FP16 inner operator: A P from its FP16 copy (a third of the
        // bytes of the FP32 hi part; a preconditioner needs no more)
        const bool from16 = c->cfg.precision == WB_PREC_FP32_IR && c->cfg.ir_inner_format == 1;
        const bool fromhm = c->cfg.precision == WB_PREC_FP32_IR && c->cfg.ir_inner_format == 2;
        if (from16) ensure_h16(c);
        auto coarse = [&] {
            launch_coarse_build(c->n, agg, nc, c->G.p, c->H.p, c->cfg.precision, c->ld,
                                from16 ? c->A16.p : nullptr,
                                from16 ? c->rs16.as<float>() : nullptr, c->alpha.as<double>(),
                                c->gmask.as<uint8_t>(), c->constrained.as<uint8_t>(),
                                c->cs_AP.as<double>(), c->vlen,
                                (from16 && c->h16_sums) || fromhm ? c->cs_S.as<double>()
                                                                  : nullptr,
                                c->cs_Ac.as<double>(), c->cs_Aci.as<double>(),
                                c->cs_Aci32.as<float>(), c->cs_WT.as<double>(),
                                c->cs_WT32.as<float>(), c->cs_status.as<int>(), bs_stream);
            HIPCHK(hipGetLastError());
        };
        bool coarse_done = false;
        if (fromhm) {
            // the coarse build starts as soon as the sketch's sums exist, beside the fill
            ensure_hm(c, [&] {
                coarse();
                coarse_done = true;
            });
        }
        if (!coarse_done) coarse();
        int st = 1;
        HIPCHK(hipMemcpyAsync(&st, c->cs_status.p, sizeof(int), hipMemcpyDeviceToHost,
                              bs_stream));
        HIPCHK(hipStreamSynchronize(bs_stream));
        if (side) {
            HIPCHK(hipEventRecord(c->ev_side, c->stream2));
            HIPCHK(hipStreamWaitEvent(c->stream, c->ev_side, 0));
        }
        // a singular or non-finite coarse inverse leaves the blocks alone
        c->cs_on = st == 0;
    }
    c->pc_assembly_ver = c->assembly_ver;
    c->pc_mask_ver = c->mask_ver;
    c->pc_cons_ver = c->cons_struct_ver;
}

// ------------------------------------------------------------------ GMRES
// deal.II SolverGMRES semantics (left preconditioning, restart after max_tmp-2 inner
// steps, SolverControl check on |gamma| with an absolute tolerance, x0 given), with the
// Arnoldi step done as classical Gram-Schmidt with one re-orthogonalization pass.
// is_float: vectors and operator in FP32 (inner solver of the refinement), else FP64.
// b and x are device vectors of the same type.  Returns 1 converged, -1 failed.
struct GmresResult {
    int status = 0, iterations = 0, restarts = 0;
    double residual = 0.0, first_residual = 0.0;
};

static GmresResult gmres(wb_ctx *c, int is_float, const void *b, void *x, double abs_tol,
                         double rel_tol, int max_steps, bool x_is_zero)
{
    const int n = c->n;
    const size_t ldv = c->vlen;
    const int m = std::max(1, c->cfg.gmres_max_tmp_vectors - 2);
    const size_t es = is_float ? 4 : 8;
    DevBuf &V = is_float ? c->V32 : c->V;
    // the cluster-order copy is only trusted for vectors this call's steps produce (the
    // basis may be reallocated below)
    c->hm_xc_for = nullptr;
    if (V.bytes < es * ldv * (m + 1)) {
        V.ensure(es * ldv * (m + 1));
        HIPCHK(hipMemsetAsync(V.p, 0, V.bytes, c->stream));
    }
    void *z = is_float ? c->z32.p : c->z.p;
    void *w = is_float ? c->w32.p : c->w.p;
    const int mode = is_float ? c->float_mode : (c->cfg.precision == WB_PREC_FP64 ? 0 : 2);
    const int parts = krylov_partials_blocks(n);
    auto vec = [&](int k) { return (void *)((char *)V.p + es * ldv * k); };
    std::vector<double> h(m + 2), gamma(m + 2), ci(m + 1), si(m + 1), Hh((size_t)(m + 1) * m),
        y(m + 1);
    GmresResult R;
    int acc = 0;
    int state = 0;
    bool first = true;
    double tol = abs_tol;
    double ratio = 1.0;  // last residual reduction factor of one step
    do {
        // v0 = P^-1 (b - A x)
        if (x_is_zero && first) {
            apply_precond(c, is_float, b, z);
        } else {
            apply_constrained(c, mode, x, w);
            launch_vec_ops(is_float ? 9 : 8, n, w, b, w, 0, c->stream);
            apply_precond(c, is_float, w, z);
        }
        launch_norm2_partials(is_float, z, n, c->Pn.as<double>(), c->stream);
        const unsigned long long s0 = ++c->hmap_seq;
        finalize_publish(c, is_float, 0, nullptr, nullptr, z, n, vec(0), s0);
        double rho = wait_hmap(c, s0)[0];
        if (first) {
            R.first_residual = rho;
            if (rel_tol > 0) tol = std::max(abs_tol, rel_tol * rho);
        }
        first = false;
        R.residual = rho;
        // SolverControl::check
        if (rho <= tol) state = 1;
        else if (std::isnan(rho) || acc >= max_steps) state = -1;
        if (state != 0) break;
        gamma[0] = rho;
        int dim = 0;
        // Arnoldi step k: v_{k+1} and column k of the Hessenberg matrix, published to the
        // host as the returned publication number.  The step does not depend on the host's
        // rotations, so step k + 1 is queued before step k's column is read whenever the
        // residual trend (rho times the last reduction ratio) stays SPEC_MARGIN above the
        // tolerance: the launches then overlap the wait.  A step queued for a check that
        // ends the cycle only writes v_{k+2} and work vectors, which nothing reads.
        auto enqueue_step = [&](int k) {
            const bool cparts = apply_constrained(c, mode, vec(k), w, true);
            apply_precond(c, is_float, w, z, cparts);
            double *h1 = c->hred.as<double>(), *h2 = h1 + KMAX;
            launch_dot_partials(is_float, V.p, ldv, k + 1, z, n, c->P1.as<double>(), c->stream);
            launch_update_dot(is_float, V.p, ldv, k + 1, nullptr, z, n, c->P2.as<double>(),
                              nullptr, 0, c->stream, 1, 0, 0, 0, nullptr, c->P1.as<double>(),
                              h1);
            launch_update_dot(is_float, V.p, ldv, k + 1, nullptr, z, n, nullptr,
                              c->Pn.as<double>(), 1, c->stream, 1, 0, 0, 0, nullptr,
                              c->P2.as<double>(), h2);
            const unsigned long long sq = ++c->hmap_seq;
            finalize_publish(c, is_float, k + 1, h1, h2, z, n, vec(k + 1), sq);
            return sq;
        };
        constexpr double SPEC_MARGIN = 8.0;
        unsigned long long queued = 0;  // publication number of step inner, if already queued
        for (int inner = 0; inner < m && state == 0; ++inner) {
            ++acc;
            const unsigned long long sk = queued ? queued : enqueue_step(inner);
            queued = 0;
            if (inner + 1 < m && rho * ratio > SPEC_MARGIN * tol) queued = enqueue_step(inner + 1);
            dim = inner + 1;
            const double *col = wait_hmap(c, sk);
            for (int i = 0; i <= dim; ++i) h[i] = col[i];
            for (int i = 0; i < inner; ++i) {
                const double s = si[i], cc = ci[i], d0 = h[i];
                h[i] = cc * d0 + s * h[i + 1];
                h[i + 1] = -s * d0 + cc * h[i + 1];
            }
            const double r = 1. / std::sqrt(h[inner] * h[inner] + h[inner + 1] * h[inner + 1]);
            si[inner] = h[inner + 1] * r;
            ci[inner] = h[inner] * r;
            h[inner] = ci[inner] * h[inner] + si[inner] * h[inner + 1];
            gamma[inner + 1] = -si[inner] * gamma[inner];
            gamma[inner] *= ci[inner];
            for (int i = 0; i < dim; ++i) Hh[(size_t)i * m + inner] = h[i];
            const double rho_new = std::fabs(gamma[dim]);
            if (rho > 0 && rho_new < rho) ratio = rho_new / rho;
            else ratio = 1.0;
            rho = rho_new;
            R.residual = rho;
            if (rho <= tol) state = 1;
            else if (std::isnan(rho) || acc >= max_steps) state = -1;
        }
        for (int i = dim - 1; i >= 0; --i) {
            double s = gamma[i];
            for (int j = i + 1; j < dim; ++j) s -= Hh[(size_t)i * m + j] * y[j];
            y[i] = s / Hh[(size_t)i * m + i];
        }
        if (dim > 0) {
            HIPCHK(hipMemcpyAsync(c->hbuf.p, y.data(), sizeof(double) * dim,
                                  hipMemcpyHostToDevice, c->stream));
            launch_axpy_basis(is_float, V.p, ldv, dim, c->hbuf.as<double>(), x, is_float, n,
                              c->stream);
        }
        if (state == 0) R.restarts++;
    } while (state == 0);
    HIPCHK(hipStreamSynchronize(c->stream));
    R.iterations = acc;
    R.status = state;
    return R;
}

// ------------------------------------------------------------------ solve_system
// Absolute tolerance of one inner FP32 solve of the refinement loop, given the current
// preconditioned outer residual rho.
static double ir_inner_target(const wb_ctx *c, double rho)
{
    const double red =
        c->float_mode >= 3 ? c->cfg.ir_inner16_reduction : c->cfg.ir_inner_reduction;
    return std::max(red * rho, 0.5 * c->cfg.solver_tolerance);
}

// The FP16 inner operator for the current assembly and masks; with a coarse space
// configured, also the sums of its rows over the coarse aggregates (cs_S) where the build
// kernel holds whole rows in registers.
static void ensure_h16(wb_ctx *c)
{
    int agg = 0, nc = 0;
    if (c->cfg.precond_coarse > 0) coarse_shape(c, agg, nc);
    if (c->h16_assembly_ver == c->assembly_ver && c->h16_mask_ver == c->mask_ver &&
        c->h16_agg == agg)
        return;
    const size_t entries = (size_t)c->n * c->ld;
    c->A16.ensure(entries * 2);
    c->rs16.ensure(sizeof(float) * c->vlen);
    if (agg > 0) c->cs_S.ensure(sizeof(double) * (size_t)nc * c->vlen);
    {
        KernelScope ks(c, WB_K_INNER, c->n);
        c->h16_sums = launch_build_h16(c->G.as<float>(), c->H.as<float>(), c->ld, c->n,
                                       c->segs.as<int>(), c->n_segs, c->gmask.as<uint8_t>(),
                                       c->A16.p, c->rs16.as<float>(), agg, nc,
                                       agg > 0 ? c->cs_S.as<double>() : nullptr, c->vlen,
                                       c->stream);
        HIPCHK(hipGetLastError());
    }
    c->h16_assembly_ver = c->assembly_ver;
    c->h16_mask_ver = c->mask_ver;
    c->h16_agg = agg;
}

// The compressed copy of the FP32 vmult operator for the current assembly and masks.
// after_ranks (optional) runs once the sketch, its aggregate sums and the rank test are
// complete on the device (the host has synchronised), before the fill is launched
static void ensure_hm(wb_ctx *c, const std::function<void()> &after_ranks)
{
    int agg = 0, nc = 0;
    if (c->cfg.precond_coarse > 0) coarse_shape(c, agg, nc);
    if (c->hm_assembly_ver == c->assembly_ver && c->hm_mask_ver == c->mask_ver &&
        c->hm_agg == agg)
        return;
    require(c->hmp.ncl > 0, WB_E_STATE, "no cluster plan (FP32_IR contexts plan at wb_set_mesh)");
    const size_t ncl = (size_t)c->hmp.ncl, nb = ncl * ncl;
    c->hm_rank.ensure(nb);
    c->hm_scale.ensure(sizeof(float) * nb);
    c->hm_Y.ensure(sizeof(float) * (size_t)c->n * ncl * HM_SK);  // [J][row][k]
    c->hm_Mx.ensure(sizeof(float) * (size_t)c->n * ncl);
    c->hm_Q.ensure(2 * nb * HM_RMAX * HM_CL);
    c->hm_B.ensure(2 * nb * HM_RMAX * HM_CL);
    // the product's x in cluster order; zero at the padding positions (p >= |J|), which
    // only the gather writes
    c->hm_T.ensure(sizeof(float) * ncl * HM_CL);
    HIPCHK(hipMemsetAsync(c->hm_T.p, 0, sizeof(float) * ncl * HM_CL, c->stream));
    c->hm_xc_for = nullptr;
    c->hm_part.ensure(sizeof(float) * c->vlen * hm_chunks(c->hmp.ncl));
    c->hm_doff.ensure(sizeof(int) * nb);
    if (agg > 0) c->cs_S.ensure(sizeof(double) * HM_SKETCH_PARTS * (size_t)nc * c->vlen);
    {
        KernelScope ks(c, WB_K_INNER, c->n);
        launch_hm_rank(c->G.as<float>(), c->H.as<float>(), c->ld, c->n, c->gmask.as<uint8_t>(),
                       c->hm_idx.as<int>(), c->hm_size.as<int>(), c->hm_seg.as<int>(),
                       c->hmp.ncl, c->hm_near.as<uint8_t>(), (float)c->cfg.ir_inner_lr_tol,
                       c->hm_Y.as<float>(), c->hm_Mx.as<float>(), c->hm_rank.as<uint8_t>(),
                       c->hm_scale.as<float>(), c->hm_Q.p, agg, nc,
                       agg > 0 ? c->cs_S.as<double>() : nullptr, c->vlen, c->stream);
        HIPCHK(hipGetLastError());
        // storage for the dense blocks only: the ranks decide which blocks are dense
        std::vector<uint8_t> rk(nb);
        HIPCHK(hipMemcpyAsync(rk.data(), c->hm_rank.p, nb, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (after_ranks) after_ranks();
        std::vector<int> doff(nb);
        int nd = 0;
        for (size_t b = 0; b < nb; ++b) doff[b] = rk[b] == HM_DENSE ? nd++ : -1;
        upload(c->hm_doff, doff.data(), nb, c->stream);
        c->hm_D.ensure(2 * (size_t)std::max(nd, 1) * HM_CL * HM_CL);
        launch_hm_fill(c->G.as<float>(), c->H.as<float>(), c->ld, c->gmask.as<uint8_t>(),
                       c->hm_idx.as<int>(), c->hm_size.as<int>(), c->hmp.ncl,
                       c->hm_rank.as<uint8_t>(), c->hm_scale.as<float>(), c->hm_Q.p, c->hm_B.p,
                       c->hm_D.p, c->hm_doff.as<int>(), c->stream);
        HIPCHK(hipGetLastError());
        // the product's work lists, planned while the fill runs
        std::vector<int> iptr, items;
        hm_plan_items(c->hmp.ncl, rk.data(), c->hmp.cl_size.data(), iptr, items);
        upload(c->hm_iptr, iptr.data(), iptr.size(), c->stream);
        upload(c->hm_items, items.data(), items.size(), c->stream);
    }
    c->hm_assembly_ver = c->assembly_ver;
    c->hm_mask_ver = c->mask_ver;
    c->hm_agg = agg;
}

// selects the operator of the FP32-vector solves for the lifetime of the guard
struct FloatMode {
    wb_ctx *c;
    int old;
    FloatMode(wb_ctx *c_, int m) : c(c_), old(c_->float_mode) { c->float_mode = m; }
    ~FloatMode() { c->float_mode = old; }
};

// b (device, internal order, vlen) must already hold the distributed right-hand side;
// the solution lands in c->vx.
static void run_solver(wb_ctx *c, wb_solve_stats *st)
{
    ensure_work(c);
    const int n = c->n;
    double t0 = now_ms();
    ensure_precond(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    double t1 = now_ms();
    if (st) st->t_precond_ms = t1 - t0;
    launch_vec_ops(6, n, nullptr, nullptr, c->vx.p, 0, c->stream);
    wb_solve_stats s{};
    if (c->cfg.precision == WB_PREC_FP64) {
        GmresResult R = gmres(c, 0, c->vb.p, c->vx.p, c->cfg.solver_tolerance, 0.0,
                              c->cfg.solver_max_steps, true);
        s.iterations = R.iterations;
        s.restarts = R.restarts;
        s.residual = R.residual;
        if (R.status != 1) {
            if (st) *st = s;
            throw WbError(WB_E_NOCONV, "GMRES did not converge (SolverControl::NoConvergence)");
        }
    } else {
        // FP64 iterative refinement around an FP32-vector GMRES on the FP32 (hi) operator,
        // its row-scaled FP16 copy or its compressed copy; the transposed system (adjoint)
        // runs its inner solves on the transposed hi part
        const int fmode = c->op_transpose ? 1
                        : (c->cfg.ir_inner_format == 2 ? 4
                                                        : (c->cfg.ir_inner_format == 1 ? 3 : 1));
        if (fmode == 3) ensure_h16(c);
        if (fmode == 4) ensure_hm(c);
        FloatMode fm(c, fmode);
        double rho_prev = 0.0;
        const size_t vb4 = sizeof(float) * c->vlen;
        for (DevBuf *b : {&c->z32, &c->w32, &c->r32, &c->d32}) {
            if (b->bytes < vb4) {
                b->ensure(vb4);
                HIPCHK(hipMemsetAsync(b->p, 0, vb4, c->stream));
            }
        }
        // r = b (x = 0)
        HIPCHK(hipMemcpyAsync(c->vr.p, c->vb.p, sizeof(double) * n, hipMemcpyDeviceToDevice,
                              c->stream));
        int total_inner = 0, sweeps = 0;
        double rho = 0.0;
        bool converged = false;
        for (;;) {
            apply_precond(c, 0, c->vr.p, c->vt.p);
            rho = device_norm(c, c->vt.p, 0);
            s.residual = rho;
            if (rho <= c->cfg.solver_tolerance) {
                converged = true;
                break;
            }
            if (std::isnan(rho) || sweeps >= c->cfg.ir_max_sweeps ||
                total_inner >= c->cfg.solver_max_steps)
                break;
            // an FP16 sweep that no longer halves the residual hands over to the FP32
            // operator for the remaining sweeps
            if (c->float_mode >= 3 && sweeps > 0 && !(rho < 0.5 * rho_prev)) c->float_mode = 1;
            rho_prev = rho;
            // inner: A32 d = r from d = 0, down to the larger of ir_inner_reduction * rho and
            // what the outer tolerance still asks for (half of it, leaving room for the
            // FP32 operator's own error in the update)
            launch_vec_ops(1, n, c->vr.p, nullptr, c->r32.p, 0, c->stream);
            launch_vec_ops(7, n, nullptr, nullptr, c->d32.p, 0, c->stream);
            GmresResult R = gmres(c, 1, c->r32.p, c->d32.p, ir_inner_target(c, rho), 0.0,
                                  c->cfg.solver_max_steps - total_inner, true);
            if (sweeps < 8) {
                s.sweep_iterations[sweeps] = R.iterations;
                s.sweep_residual[sweeps] = rho;
            }
            total_inner += R.iterations;
            s.restarts += R.restarts;
            launch_vec_ops(2, n, c->d32.p, nullptr, c->vx.p, 0, c->stream);
            // r = b - A x  (FP64-accurate operator: hi + lo)
            apply_constrained(c, 2, c->vx.p, c->w.p);
            launch_vec_ops(8, n, c->w.p, c->vb.p, c->vr.p, 0, c->stream);
            sweeps++;
        }
        s.iterations = total_inner;
        s.ir_sweeps = sweeps;
        if (!converged) {
            if (st) *st = s;
            throw WbError(WB_E_NOCONV, "iterative refinement did not converge");
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    double t2 = now_ms();
    s.t_precond_ms = t1 - t0;
    s.t_gmres_ms = t2 - t1;
    // true relative residual of the constrained system with the FP64-accurate operator
    // (the refinement loop already holds it in vr)
    {
        double rn;
        if (c->cfg.precision == WB_PREC_FP64) {
            apply_constrained(c, 0, c->vx.p, c->w.p);
            launch_vec_ops(8, n, c->w.p, c->vb.p, c->vt.p, 0, c->stream);
            rn = device_norm(c, c->vt.p, 0);
        } else {
            rn = device_norm(c, c->vr.p, 0);
        }
        const double bn = device_norm(c, c->vb.p, 0);
        s.true_rel_residual = bn > 0 ? rn / bn : rn;
    }
    if (st) {
        s.t_alpha_ms = st->t_alpha_ms;
        s.t_rhs_ms = st->t_rhs_ms;
        *st = s;
    }
}

// rhs = compute_rhs(tmp) then distribute_rhs (constrained rows = inhomogeneity), in vb
static void build_rhs(wb_ctx *c, const void *tmp_dev)
{
    gemv(c, c->cfg.precision == WB_PREC_FP64 ? 0 : 2, 1, tmp_dev, tmp_dev, c->vb.p, true);
    launch_constraint_rows(0, c->n_lines, c->con_rows.as<int>(), c->con_ptr.as<int>(),
                           c->con_cols.as<int>(), c->con_w.as<double>(), c->con_inhom.as<double>(),
                           1, nullptr, c->vb.p, c->stream);
}

// L2 projections at the current support points (caller order, n x 3 each, either output
// may be NULL): the cell normals, normalised per node when `normalize`, and the surface
// gradient of u (caller order; NULL reads as 0).  Jacobi-preconditioned CG on the Q1 mass
// matrix to a relative residual of 1e-14 per column, in place of the reference's direct
// solve (the mass matrix is SPD and, diagonally scaled, well conditioned).
static void project_fields(wb_ctx *c, const double *u, int normalize, double *normals,
                           double *grads, const double *psi = nullptr)
{
    const int n = c->n, nc = c->plan.nc, nq = c->quad.q_order * c->quad.q_order;
    hipStream_t s = c->stream;
    const size_t v6 = sizeof(double) * 6 * (size_t)n;
    c->pj_emass.ensure(sizeof(double) * 10 * (size_t)nc);
    c->pj_erhs.ensure(sizeof(double) * 24 * (size_t)nc);
    for (DevBuf *b : {&c->pj_B, &c->pj_X, &c->pj_R, &c->pj_Z, &c->pj_P, &c->pj_Q}) b->ensure(v6);
    c->pj_dinv.ensure(sizeof(double) * n);
    const int nb = proj_blocks(n);
    c->pj_part.ensure(sizeof(double) * 6 * nb);
    c->pj_part2.ensure(sizeof(double) * 12 * nb);
    c->pj_state.ensure(sizeof(double) * 24);
    const double *ud = nullptr;
    if (u) {
        std::vector<double> ui;
        to_internal(c, u, ui);
        upload(c->pj_u, ui.data(), ui.size(), s);
        ud = c->pj_u.as<double>();
    }
    const double *pd = nullptr;
    if (psi) {
        std::vector<double> pi;
        to_internal(c, psi, pi);
        upload(c->pj_psi, pi.data(), pi.size(), s);
        pd = c->pj_psi.as<double>();
    }
    double *X = c->pj_X.as<double>(), *R = c->pj_R.as<double>(), *Z = c->pj_Z.as<double>(),
           *P = c->pj_P.as<double>(), *Q = c->pj_Q.as<double>(), *st = c->pj_state.as<double>();
    launch_proj_setup(c->X.as<double>(), c->cell_int.as<int>(), c->orient.as<int8_t>(), ud, pd,
                      nc, nq, c->qpts.as<double>(), c->qw.as<double>(), c->inc_ptr.as<int>(),
                      c->inc_item.as<int>(), n, c->pj_emass.as<double>(),
                      c->pj_erhs.as<double>(), c->pj_B.as<double>(), c->pj_dinv.as<double>(), s);
    launch_proj_cg_start(c->pj_B.as<double>(), c->pj_dinv.as<double>(), n, X, R, Z, P,
                         c->pj_part2.as<double>(), st, s);
    HIPCHK(hipGetLastError());
    const double tol = 1e-14;
    const int max_it = 1000, batch = 8;
    double hs[24];
    int it = 0;
    for (;;) {
        HIPCHK(hipMemcpyAsync(hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        bool done = true;
        for (int e = 0; e < 6; ++e) done = done && hs[18 + e] <= tol * tol * hs[12 + e];
        if (done) break;
        require(it < max_it, WB_E_NOCONV, "L2 projection: CG on the mass matrix did not converge");
        for (int k = 0; k < batch; ++k, ++it)
            launch_proj_cg_iter(c->inc_ptr.as<int>(), c->inc_item.as<int>(),
                                c->cell_int.as<int>(), c->pj_emass.as<double>(),
                                c->pj_dinv.as<double>(), n, it, X, R, Z, P, Q,
                                c->pj_part.as<double>(), c->pj_part2.as<double>(), st, s);
        HIPCHK(hipGetLastError());
    }
    c->proj_iterations = it;
    std::vector<double> xi(6 * (size_t)n);
    HIPCHK(hipMemcpyAsync(xi.data(), X, v6, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int d = 0; d < n; ++d) {
        const double *x = xi.data() + 6 * (size_t)c->plan.perm[d];
        if (normals) {
            double sc = 1.0;
            if (normalize) {
                const double l = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
                require(l > 0.0, WB_E_ARG, "zero projected normal");
                sc = 1.0 / l;
            }
            for (int e = 0; e < 3; ++e) normals[3 * (size_t)d + e] = x[e] * sc;
        }
        if (grads)
            for (int e = 0; e < 3; ++e) grads[3 * (size_t)d + e] = x[3 + e];
    }
}

// Whether two Dirichlet dofs share a double-node set: the only case in which the
// constraint rules read nodal normals and surface gradients.
static bool has_dirichlet_pairs(const wb_ctx *c)
{
    const MeshPlan &p = c->plan;
    for (int i = 0; i < c->n; ++i) {
        if (c->surface[i] != 1.0) continue;
        for (int q = p.dn_ptr[i]; q < p.dn_ptr[i + 1]; ++q)
            if (p.dn_idx[q] != i && c->surface[p.dn_idx[q]] == 1.0) return true;
    }
    return false;
}

// BEMProblem::compute_constraints (source/bem_problem.cc:990-1105) for tmp_rhs.  The
// reference starts with comp_dom.compute_normals() and compute_surface_gradients(tmp_rhs)
// (:996-997); here they run on the device, and only when the rules will read them.
static void auto_constraints(wb_ctx *c, const double *tmp_rhs, ConstraintSet &cs)
{
    const double *nn = c->cons_normals.empty() ? nullptr : c->cons_normals.data();
    const double *gg = c->cons_grads.empty() ? nullptr : c->cons_grads.data();
    std::vector<double> dn, dg;
    if (c->cons_project && (!nn || !gg) && has_dirichlet_pairs(c)) {
        std::vector<double> phi(c->n);
        for (int i = 0; i < c->n; ++i) phi[i] = tmp_rhs[i] * c->surface[i];
        dn.resize(3 * (size_t)c->n);
        dg.resize(3 * (size_t)c->n);
        project_fields(c, phi.data(), 1, nn ? nullptr : dn.data(), gg ? nullptr : dg.data());
        if (!nn) nn = dn.data();
        if (!gg) gg = dg.data();
    }
    std::string e = build_constraints(c->n, c->plan.dn_ptr.data(), c->plan.dn_idx.data(),
                                      c->surface.data(), nn, gg, tmp_rhs, cs);
    require(e.empty(), WB_E_UNSUPPORTED, e.c_str());
}

static void refresh_auto_constraints(wb_ctx *c, const double *tmp_rhs_caller)
{
    if (!c->cons_auto) return;
    ConstraintSet cs;
    auto_constraints(c, tmp_rhs_caller, cs);
    install_constraints(c, cs);
}

static void check_ready(wb_ctx *c, bool need_assembly)
{
    require(c->have_mesh, WB_E_STATE, "wb_set_mesh has not been called");
    require(c->have_masks, WB_E_STATE, "wb_set_masks has not been called");
    if (need_assembly) require(c->assembled, WB_E_STATE, "wb_assemble has not been called");
}

static void solve_system_dev(wb_ctx *c, const double *tmp_dev_internal, wb_solve_stats *st)
{
    double t0 = now_ms();
    build_rhs(c, tmp_dev_internal);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (st) {
        st->t_alpha_ms = 0.0;  // alpha comes out of the assembly kernels
        st->t_rhs_ms = now_ms() - t0;
    }
    run_solver(c, st);
}

// ------------------------------------------------------------------ several right-hand sides
// Y[:, v] = ConstrainedOperator(A) X[:, v] for the k columns (column strides ldx / ldy).
// chunk partials of the batched product for k vectors over the unconstrained rows
static double *multi_part(wb_ctx *c, int mode, int k)
{
    const size_t need =
        sizeof(double) * (size_t)multi_chunks(mode, c->n_free, (c->n + 3) / 4) * k * c->n_free;
    c->mw.part.ensure(std::max<size_t>(need, 8));
    return c->mw.part.as<double>();
}

static void apply_constrained_multi(wb_ctx *c, int mode, int k, const void *X, size_t ldx,
                                    void *Y, size_t ldy)
{
    require(!c->pure_neumann, WB_E_UNSUPPORTED,
            "several right-hand sides with all-Neumann masks are not supported");
    if (mode == 4) {
        // the compressed operator, read once per group of up to 8 vectors, with the
        // constraint lines in its last kernel
        KernelScope ks(c, WB_K_GEMVC, (long long)c->n * k);
        size_t xcf = 0;
        const size_t pf = hm_multi_work_floats(c->hmp.ncl, c->vlen, &xcf);
        c->mw.hxc.ensure(sizeof(float) * xcf);
        c->mw.hpart.ensure(sizeof(float) * pf);
        launch_hm_gemv_multi(c->hm_idx.as<int>(), c->hm_size.as<int>(), c->hmp.ncl,
                             c->hm_rank.as<uint8_t>(), c->hm_scale.as<float>(), c->hm_Q.p,
                             c->hm_B.p, c->hm_D.p, c->hm_doff.as<int>(), c->hm_iptr.as<int>(),
                             c->hm_items.as<int>(), c->n, k, (const float *)X, ldx,
                             c->diag_op.as<double>(), c->mw.hxc.as<float>(),
                             c->mw.hpart.as<float>(), c->vlen, (float *)Y, ldy,
                             c->n_lines > 0 ? c->con_line.as<int>() : nullptr,
                             c->con_ptr.as<int>(), c->con_cols.as<int>(), c->con_w.as<double>(),
                             c->stream);
        return;
    }
    const int is_float = mode == 1 || mode == 3;
    const size_t es = is_float ? 4 : 8;
    {
        KernelScope ks(c,
                       mode == 0 ? WB_K_GEMV64
                                 : (mode == 1 ? WB_K_GEMV32 : (mode == 3 ? WB_K_GEMV16 : WB_K_GEMVHL)),
                       (long long)c->n_free);
        double *part = multi_part(c, mode, k);
        // mode 3 reads the FP16 operator (A16) and its row scales in place of G and H, and
        // takes the vector-scale and split-vector workspace in place of Glo
        const void *Gm = mode == 3 ? c->A16.p : c->G.p;
        const void *Hm = mode == 3 ? c->rs16.p : c->H.p;
        if (mode == 3) c->mw.xs.ensure(256 + 4 * (size_t)k * c->ld);
        const void *Glm = mode == 3 ? c->mw.xs.p : c->Glo.p;
        launch_gemv_sel_multi(mode, Gm, Hm, Glm, c->Hlo.p, c->ld, c->segs.as<int>(),
                              c->n_segs, c->gmask.as<uint8_t>(), 0, 1.0, -1.0, k, X, ldx,
                              c->diag_op.as<double>(), X, c->free_rows.as<int>(), c->n_free, Y,
                              ldy, part, (c->n + 3) / 4, c->stream);
    }
    for (int v = 0; v < k; ++v)
        launch_constraint_rows(is_float, c->n_lines, c->con_rows.as<int>(), c->con_ptr.as<int>(),
                               c->con_cols.as<int>(), c->con_w.as<double>(),
                               c->con_inhom.as<double>(), 0, (const char *)X + es * ldx * v,
                               (char *)Y + es * ldy * v, c->stream);
}

static void precond_multi(wb_ctx *c, int is_float, int k, const void *X, void *Y)
{
    if (c->cfg.preconditioner == WB_PRECOND_BLOCK_JACOBI)
        block_precond(c, is_float, k, X, Y);
    else
        HIPCHK(hipMemcpyAsync(Y, X, (is_float ? 4 : 8) * c->vlen * k, hipMemcpyDeviceToDevice,
                              c->stream));
}

// k GMRES instances in lockstep (same restart cycle), deal.II semantics per instance.
// B, X: k x vlen (column stride vlen).  state[v] in/out: 0 iterate, 1 converged, -1 failed;
// instances entering with state != 0 are left alone.  Returns per-instance iterations.
static std::vector<int> gmres_multi(wb_ctx *c, int is_float, int k, const void *B, void *X,
                                    const std::vector<double> &abs_tol, int max_steps,
                                    std::vector<int> &state, std::vector<double> &resid)
{
    auto &M = c->mw;
    const int n = c->n;
    const size_t vl = c->vlen;
    const int m = std::max(1, c->cfg.gmres_max_tmp_vectors - 2);
    const size_t es = is_float ? 4 : 8;
    const size_t sV = (size_t)(m + 1) * vl;
    DevBuf &V = is_float ? M.V32 : M.V;
    if (V.bytes < es * sV * k) {
        V.ensure(es * sV * k);
        HIPCHK(hipMemsetAsync(V.p, 0, V.bytes, c->stream));
    }
    void *Z = is_float ? M.Z32.p : M.Z.p;
    void *W = is_float ? M.W32.p : M.W.p;
    const int mode = is_float ? c->float_mode : (c->cfg.precision == WB_PREC_FP64 ? 0 : 2);
    const int parts = krylov_partials_blocks(n);
    const size_t sH = 132;
    auto basis = [&](int i) { return (void *)((char *)V.p + es * vl * i); };
    std::vector<std::vector<double>> gamma(k, std::vector<double>(m + 2)),
        ci(k, std::vector<double>(m + 1)), si(k, std::vector<double>(m + 1)),
        Hh(k, std::vector<double>((size_t)(m + 1) * m));
    std::vector<double> h(m + 2), y(m + 1);
    std::vector<int> its(k, 0), dim(k, 0);
    int *act = M.apin;
    const std::vector<double> &tol = abs_tol;
    bool first = true;
    auto upload_act = [&]() {
        HIPCHK(hipStreamSynchronize(c->stream));  // previous copy out of apin has finished
        for (int v = 0; v < k; ++v) act[v] = state[v] == 0;
        HIPCHK(hipMemcpyAsync(M.act.p, act, sizeof(int) * k, hipMemcpyHostToDevice, c->stream));
    };
    for (;;) {
        upload_act();
        if (first) {
            precond_multi(c, is_float, k, B, Z);
        } else {
            apply_constrained_multi(c, mode, k, X, vl, W, vl);
            launch_vec_ops(is_float ? 9 : 8, (int)(vl * k), W, B, W, 0, c->stream);
            precond_multi(c, is_float, k, W, Z);
        }
        launch_norm2_partials(is_float, Z, n, M.Pn.as<double>(), c->stream, k, vl,
                              M.act.as<int>());
        launch_finalize(is_float, 0, nullptr, nullptr, M.Pn.as<double>(), Z, n, basis(0),
                        M.H.as<double>(), c->stream, k, vl, sV, 0, sH, M.act.as<int>());
        HIPCHK(hipMemcpyAsync(M.hpin, M.H.p, sizeof(double) * sH * k, hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        bool any = false;
        for (int v = 0; v < k; ++v) {
            if (state[v] != 0) continue;
            const double rho = M.hpin[v * sH];
            resid[v] = rho;
            if (rho <= tol[v]) state[v] = 1;
            else if (std::isnan(rho) || its[v] >= max_steps) state[v] = -1;
            else {
                gamma[v][0] = rho;
                dim[v] = 0;
                any = true;
            }
        }
        first = false;
        if (!any) break;
        upload_act();
        for (int inner = 0; inner < m; ++inner) {
            bool still = false;
            for (int v = 0; v < k; ++v) still |= (act[v] != 0);
            if (!still) break;
            apply_constrained_multi(c, mode, k, basis(inner), sV, W, vl);
            precond_multi(c, is_float, k, W, Z);
            const int kc = inner + 1;
            const size_t sR = 2 * KMAX;
            double *h1 = M.hred.as<double>(), *h2 = h1 + KMAX;
            launch_dot_partials(is_float, V.p, vl, kc, Z, n, M.P1.as<double>(), c->stream, k, sV,
                                vl, M.act.as<int>());
            launch_reduce_parts(kc, M.P1.as<double>(), n, h1, c->stream, k, sR, M.act.as<int>());
            launch_update_dot(is_float, V.p, vl, kc, h1, Z, n, M.P2.as<double>(), nullptr, 0,
                              c->stream, k, sV, vl, sR, M.act.as<int>());
            launch_reduce_parts(kc, M.P2.as<double>(), n, h2, c->stream, k, sR, M.act.as<int>());
            launch_update_dot(is_float, V.p, vl, kc, h2, Z, n, nullptr, M.Pn.as<double>(), 1,
                              c->stream, k, sV, vl, sR, M.act.as<int>());
            launch_finalize(is_float, kc, h1, h2, M.Pn.as<double>(), Z, n, basis(kc),
                            M.H.as<double>(), c->stream, k, vl, sV, sR, sH, M.act.as<int>());
            HIPCHK(hipMemcpyAsync(M.hpin, M.H.p, sizeof(double) * sH * k, hipMemcpyDeviceToHost,
                                  c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            bool changed = false;
            for (int v = 0; v < k; ++v) {
                if (!act[v]) continue;
                ++its[v];
                dim[v] = kc;
                for (int i = 0; i <= kc; ++i) h[i] = M.hpin[v * sH + i];
                for (int i = 0; i < inner; ++i) {
                    const double s = si[v][i], cc = ci[v][i], d0 = h[i];
                    h[i] = cc * d0 + s * h[i + 1];
                    h[i + 1] = -s * d0 + cc * h[i + 1];
                }
                const double r = 1. / std::sqrt(h[inner] * h[inner] + h[inner + 1] * h[inner + 1]);
                si[v][inner] = h[inner + 1] * r;
                ci[v][inner] = h[inner] * r;
                h[inner] = ci[v][inner] * h[inner] + si[v][inner] * h[inner + 1];
                gamma[v][inner + 1] = -si[v][inner] * gamma[v][inner];
                gamma[v][inner] *= ci[v][inner];
                for (int i = 0; i < kc; ++i) Hh[v][(size_t)i * m + inner] = h[i];
                const double rho = std::fabs(gamma[v][kc]);
                resid[v] = rho;
                if (rho <= tol[v]) state[v] = 1;
                else if (std::isnan(rho) || its[v] >= max_steps) state[v] = -1;
                if (state[v] != 0) {
                    act[v] = 0;
                    changed = true;
                }
            }
            if (changed)
                HIPCHK(hipMemcpyAsync(M.act.p, act, sizeof(int) * k, hipMemcpyHostToDevice,
                                      c->stream));
        }
        // x_v += V_v y_v for every instance that ran this cycle (y_v staged at H + v*sH)
        for (int v = 0; v < k; ++v) {
            const int d = dim[v];
            if (d == 0) continue;
            for (int i = d - 1; i >= 0; --i) {
                double s = gamma[v][i];
                for (int j = i + 1; j < d; ++j) s -= Hh[v][(size_t)i * m + j] * y[j];
                y[i] = s / Hh[v][(size_t)i * m + i];
            }
            std::copy(y.begin(), y.begin() + d, M.hpin + v * sH);
        }
        HIPCHK(hipMemcpyAsync(M.H.p, M.hpin, sizeof(double) * sH * k, hipMemcpyHostToDevice,
                              c->stream));
        for (int v = 0; v < k; ++v) {
            if (dim[v] == 0) continue;
            launch_axpy_basis(is_float, (const char *)V.p + es * sV * v, vl, dim[v],
                              M.H.as<double>() + v * sH, (char *)X + es * vl * v, is_float, n,
                              c->stream);
            dim[v] = 0;
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return its;
}

static void solve_multi(wb_ctx *c, int k, const double *tmp_rhs, double *phi, double *dphi,
                        const double *inhom_k, wb_solve_stats *st)
{
    require(!c->pure_neumann, WB_E_UNSUPPORTED,
            "several right-hand sides with all-Neumann masks are not supported");
    auto &M = c->mw;
    const int n = c->n;
    const size_t vl = c->vlen;
    ensure_work(c);
    const int parts = krylov_partials_blocks(n);
    const size_t d8 = sizeof(double) * vl * k, f4 = sizeof(float) * vl * k;
    for (DevBuf *b : {&M.Z, &M.W, &M.X, &M.B, &M.R, &M.T})
        if (b->bytes < d8) {
            b->ensure(d8);
            HIPCHK(hipMemsetAsync(b->p, 0, d8, c->stream));
        }
    for (DevBuf *b : {&M.Z32, &M.W32, &M.D32, &M.R32})
        if (b->bytes < f4) {
            b->ensure(f4);
            HIPCHK(hipMemsetAsync(b->p, 0, f4, c->stream));
        }
    M.P1.ensure(sizeof(double) * parts * KMAX * k);
    M.P2.ensure(sizeof(double) * parts * KMAX * k);
    M.Pn.ensure(sizeof(double) * parts * k);
    M.H.ensure(sizeof(double) * 132 * k);
    M.hred.ensure(sizeof(double) * 2 * KMAX * k);
    M.act.ensure(sizeof(int) * k);
    if (M.k < k) {
        if (M.hpin) (void)hipHostFree(M.hpin);
        if (M.apin) (void)hipHostFree(M.apin);
        M.hpin = nullptr;
        M.apin = nullptr;
        M.k = 0;
        HIPCHK(hipHostMalloc((void **)&M.hpin, sizeof(double) * 132 * k, 0));
        HIPCHK(hipHostMalloc((void **)&M.apin, sizeof(int) * k, 0));
        M.k = k;
    }
    // per right-hand side constraint inhomogeneities (the line structure is shared)
    const int nl = c->n_lines;
    std::vector<double> inh((size_t)std::max(nl, 1) * k, 0.0);
    for (int v = 0; v < k; ++v) {
        const double *t = tmp_rhs + (size_t)v * n;
        if (c->cons_auto) {
            ConstraintSet cs;
            auto_constraints(c, t, cs);
            if (v == 0) install_constraints(c, cs);
            require(cs.rows == c->cons.rows && cs.ptr == c->cons.ptr &&
                        cs.cols == c->cons.cols && cs.w == c->cons.w,
                    WB_E_STATE, "constraint structure differs between right-hand sides");
            std::copy(cs.inhom.begin(), cs.inhom.end(), inh.begin() + (size_t)v * nl);
        } else if (inhom_k) {
            std::copy(inhom_k + (size_t)v * nl, inhom_k + (size_t)(v + 1) * nl,
                      inh.begin() + (size_t)v * nl);
        } else {
            std::copy(c->cons.inhom.begin(), c->cons.inhom.end(), inh.begin() + (size_t)v * nl);
        }
    }
    upload(M.inh, inh.data(), inh.size(), c->stream);
    // tmp_rhs -> T (internal order), B = distributed rhs per column
    std::vector<double> tmp;
    for (int v = 0; v < k; ++v) {
        to_internal(c, tmp_rhs + (size_t)v * n, tmp);
        HIPCHK(hipMemcpyAsync((double *)M.T.p + vl * v, tmp.data(), sizeof(double) * vl,
                              hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    const int mode64 = c->cfg.precision == WB_PREC_FP64 ? 0 : 2;
    double t0 = now_ms();
    {
        KernelScope ks(c, mode64 == 0 ? WB_K_GEMV64 : WB_K_GEMVHL, (long long)c->n_free);
        double *part = multi_part(c, mode64, k);
        launch_gemv_sel_multi(mode64, c->G.p, c->H.p, c->Glo.p, c->Hlo.p, c->ld,
                              c->segs.as<int>(), c->n_segs, c->gmask.as<uint8_t>(), 1, -1.0, 1.0,
                              k, M.T.p, vl, c->diag_rhs.as<double>(), M.T.p,
                              c->free_rows.as<int>(), c->n_free, M.B.p, vl, part,
                              (c->n + 3) / 4, c->stream);
    }
    for (int v = 0; v < k; ++v)
        launch_constraint_rows(0, nl, c->con_rows.as<int>(), c->con_ptr.as<int>(),
                               c->con_cols.as<int>(), c->con_w.as<double>(),
                               M.inh.as<double>() + (size_t)v * nl, 1, nullptr,
                               (double *)M.B.p + vl * v, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    double t1 = now_ms();
    ensure_precond(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    double t2 = now_ms();
    launch_vec_ops(6, (int)(vl * k), nullptr, nullptr, M.X.p, 0, c->stream);
    std::vector<int> state(k, 0);
    std::vector<double> resid(k, 0.0);
    wb_solve_stats s{};
    if (c->cfg.precision == WB_PREC_FP64) {
        std::vector<int> its =
            gmres_multi(c, 0, k, M.B.p, M.X.p, std::vector<double>(k, c->cfg.solver_tolerance),
                        c->cfg.solver_max_steps, state, resid);
        s.iterations = *std::max_element(its.begin(), its.end());
    } else {
        HIPCHK(hipMemcpyAsync(M.R.p, M.B.p, d8, hipMemcpyDeviceToDevice, c->stream));
        int sweeps = 0, total = 0;
        std::vector<int> done(k, 0);
        // inner solves on the compressed operator (ir_inner_format 2), the FP16 one (1) or
        // the FP32 one (0)
        const int fmode = c->cfg.ir_inner_format == 2 ? 4 : (c->cfg.ir_inner_format == 1 ? 3 : 1);
        if (fmode == 3) ensure_h16(c);
        if (fmode == 4) ensure_hm(c);
        FloatMode fm(c, fmode);
        std::vector<double> prev(k, 0.0);
        for (;;) {
            precond_multi(c, 0, k, M.R.p, M.T.p);
            launch_norm2_partials(0, M.T.p, n, M.Pn.as<double>(), c->stream, k, vl);
            std::vector<double> pn((size_t)parts * k);
            HIPCHK(hipMemcpyAsync(pn.data(), M.Pn.p, sizeof(double) * parts * k,
                                  hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            bool all = true, bad = false;
            for (int v = 0; v < k; ++v) {
                double q = 0.0;
                for (int p = 0; p < parts; ++p) q += pn[(size_t)v * parts + p];
                resid[v] = std::sqrt(q);
                done[v] = resid[v] <= c->cfg.solver_tolerance;
                all &= done[v] != 0;
                bad |= std::isnan(resid[v]);
            }
            if (all || bad || sweeps >= c->cfg.ir_max_sweeps ||
                total >= c->cfg.solver_max_steps)
                break;
            // an FP16 or compressed sweep that did not halve some residual hands the
            // remaining sweeps of every right-hand side to the FP32 operator
            if (c->float_mode >= 3 && sweeps > 0)
                for (int v = 0; v < k; ++v)
                    if (!done[v] && !(resid[v] < 0.5 * prev[v])) c->float_mode = 1;
            prev = resid;
            launch_vec_ops(1, (int)(vl * k), M.R.p, nullptr, M.R32.p, 0, c->stream);
            launch_vec_ops(7, (int)(vl * k), nullptr, nullptr, M.D32.p, 0, c->stream);
            std::vector<int> st_in(k);
            std::vector<double> r_in(k), t_in(k);
            for (int v = 0; v < k; ++v) {
                st_in[v] = done[v] ? 1 : 0;
                t_in[v] = ir_inner_target(c, resid[v]);
            }
            std::vector<int> its = gmres_multi(c, 1, k, M.R32.p, M.D32.p, t_in,
                                               c->cfg.solver_max_steps - total, st_in, r_in);
            total += *std::max_element(its.begin(), its.end());
            launch_vec_ops(2, (int)(vl * k), M.D32.p, nullptr, M.X.p, 0, c->stream);
            apply_constrained_multi(c, 2, k, M.X.p, vl, M.W.p, vl);
            launch_vec_ops(8, (int)(vl * k), M.W.p, M.B.p, M.R.p, 0, c->stream);
            sweeps++;
        }
        s.iterations = total;
        s.ir_sweeps = sweeps;
        for (int v = 0; v < k; ++v) state[v] = done[v] ? 1 : -1;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    double t3 = now_ms();
    for (int v = 0; v < k; ++v) s.residual = std::max(s.residual, resid[v]);
    // true relative residuals with the FP64-accurate operator (the refinement loop already
    // holds them in R: its last product used the same operator on the final X)
    const void *Rt = M.R.p;
    if (mode64 == 0) {
        apply_constrained_multi(c, mode64, k, M.X.p, vl, M.W.p, vl);
        launch_vec_ops(8, (int)(vl * k), M.W.p, M.B.p, M.T.p, 0, c->stream);
        Rt = M.T.p;
    }
    for (int v = 0; v < k; ++v) {
        const double rn = device_norm(c, (const double *)Rt + vl * v, 0);
        const double bn = device_norm(c, (double *)M.B.p + vl * v, 0);
        s.true_rel_residual = std::max(s.true_rel_residual, bn > 0 ? rn / bn : rn);
    }
    s.t_rhs_ms = t1 - t0;
    s.t_precond_ms = t2 - t1;
    s.t_gmres_ms = t3 - t2;
    if (st) *st = s;
    for (int v = 0; v < k; ++v)
        require(state[v] == 1, WB_E_NOCONV, "GMRES did not converge for every right-hand side");
    std::vector<double> xi(vl * k);
    HIPCHK(hipMemcpyAsync(xi.data(), M.X.p, d8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int v = 0; v < k; ++v)
        for (int d = 0; d < n; ++d) {
            const double val = xi[vl * v + c->plan.perm[d]];
            if (c->surface[d] == 0)
                phi[(size_t)v * n + d] = val;
            else
                dphi[(size_t)v * n + d] = val;
        }
}

// ==================================================================== C-ABI
extern "C" {

void wb_config_default(wb_config *cfg)
{
    if (!cfg) return;
    cfg->device = 0;
    cfg->quadrature_order = 4;
    cfg->singular_quadrature_order = 5;
    cfg->solver_tolerance = 1e-10;
    cfg->solver_max_steps = 1000;
    cfg->gmres_max_tmp_vectors = 100;
    cfg->precision = WB_PREC_FP64;
    cfg->preconditioner = WB_PRECOND_BLOCK_JACOBI;
    cfg->precond_block = 64;
    cfg->ir_max_sweeps = 10;
    cfg->ir_inner_reduction = 1e-7;
    cfg->assembly_tile = 16;
    cfg->ir_inner_format = 2;
    cfg->ir_inner16_reduction = 3e-4;
    cfg->precond_coarse = CS_MAX64;
    cfg->assembly_strips = 1;
    cfg->ir_inner_lr_tol = 1e-4;
}

const char *wb_version(void) { return "wavebem_amd 0.1 (gfx950)"; }

wb_status wb_create(wb_ctx **out, const wb_config *cfg)
{
    if (!out) return WB_E_ARG;
    *out = nullptr;
    wb_ctx *c = new (std::nothrow) wb_ctx();
    if (!c) return WB_E_OOM;
    if (cfg)
        c->cfg = *cfg;
    else
        wb_config_default(&c->cfg);
    const wb_config &k = c->cfg;
    if (k.quadrature_order < 1 || k.quadrature_order > MAX_NQ1 ||
        k.singular_quadrature_order < 1 || k.singular_quadrature_order > MAX_SING ||
        k.gmres_max_tmp_vectors < 3 || k.gmres_max_tmp_vectors > 129 ||
        (k.precision != WB_PREC_FP64 && k.precision != WB_PREC_FP32_IR) ||
        k.precond_block < 1 || k.precond_block > 128 || k.solver_max_steps < 0 ||
        k.assembly_tile < 1 || k.assembly_tile > 16 ||
        k.ir_inner_format < 0 || k.ir_inner_format > 2 ||
        !(k.ir_inner_lr_tol > 0 && k.ir_inner_lr_tol < 1) ||
        !(k.ir_inner16_reduction > 0 && k.ir_inner16_reduction < 1) ||
        k.precond_coarse < 0 || k.precond_coarse > CS_MAX ||
        (k.assembly_strips != 0 && k.assembly_strips != 1) ||
        !(k.solver_tolerance >= 0)) {
        delete c;
        return WB_E_ARG;
    }
    wb_status st = guarded(c, [&] {
        HIPCHK(hipSetDevice(k.device));
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&c->ev_main, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->ev_side, hipEventDisableTiming));
        // the no-allocation entry point's work space on this device
        HIPCHK(noalloc_setup(k.device));
        build_quadrature(k.quadrature_order, k.singular_quadrature_order, c->quad);
    });
    if (st != WB_OK) {
        delete c;
        return st;
    }
    c->pc_bs = k.precond_block;
    *out = c;
    return WB_OK;
}

void wb_destroy(wb_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    for (DevBuf *b : c->all_bufs()) b->release();
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->hpin) (void)hipHostFree(c->hpin);
    if (c->hmap) (void)hipHostFree(c->hmap);
    if (c->mw.hpin) (void)hipHostFree(c->mw.hpin);
    if (c->mw.apin) (void)hipHostFree(c->mw.apin);
    if (c->ev_main) (void)hipEventDestroy(c->ev_main);
    if (c->ev_side) (void)hipEventDestroy(c->ev_side);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *wb_last_error(const wb_ctx *c) { return c ? c->err.c_str() : "null context"; }

wb_status wb_set_mesh(wb_ctx *c, int n, int nc, const double *support, const int32_t *cell_dofs,
                      const int8_t *cell_orient, const int32_t *cell_material,
                      const int32_t *dn_ptr, const int32_t *dn_idx)
{
    return guarded(c, [&] {
        require(n > 0 && nc > 0 && support && cell_dofs && cell_orient, WB_E_ARG,
                "wb_set_mesh: bad sizes or null pointers");
        require((dn_ptr == nullptr) == (dn_idx == nullptr), WB_E_ARG,
                "wb_set_mesh: dn_ptr and dn_idx must both be given or both NULL");
        for (int k = 0; k < nc; ++k)
            require(cell_orient[k] == 1 || cell_orient[k] == -1, WB_E_ARG,
                    "wb_set_mesh: cell_orient must be +-1");
        HIPCHK(hipSetDevice(c->cfg.device));
        MeshPlan P;
        std::string e = plan_mesh(n, nc, support, cell_dofs, cell_material, dn_ptr, dn_idx,
                                  c->cfg.assembly_tile, c->cfg.assembly_strips, P);
        require(e.empty(), WB_E_ARG, e.c_str());
        c->have_mesh = c->have_masks = c->assembled = c->have_cons = false;
        c->plan = std::move(P);
        c->cell_dofs_h.assign(cell_dofs, cell_dofs + 4 * (size_t)nc);
        if (cell_material)
            c->cell_mat_h.assign(cell_material, cell_material + nc);
        else
            c->cell_mat_h.clear();
        c->n = n;
        c->ld = c->plan.ld;
        c->vlen = c->ld;
        // vectors and Krylov bases kept from an earlier mesh take the new layout: clear
        // them whole, so that the padding [n, vlen) of every vector reads zero
        for (DevBuf *b : {&c->z, &c->w, &c->vx, &c->vb, &c->vr, &c->vt, &c->vs, &c->V, &c->V32,
                          &c->z32, &c->w32, &c->r32, &c->d32})
            if (b->p) HIPCHK(hipMemsetAsync(b->p, 0, b->bytes, c->stream));
        for (DevBuf *b : c->mw.all())
            if (b->p) HIPCHK(hipMemsetAsync(b->p, 0, b->bytes, c->stream));
        const MeshPlan &p = c->plan;
        std::vector<double> Xi(3 * (size_t)n);
        for (int d = 0; d < n; ++d)
            for (int k = 0; k < 3; ++k) Xi[3 * p.perm[d] + k] = support[3 * d + k];
        hipStream_t s = c->stream;
        upload(c->X, Xi.data(), Xi.size(), s);
        upload(c->cell_int, p.cell_int.data(), p.cell_int.size(), s);
        upload(c->orient, cell_orient, nc, s);
        upload(c->cell_cls, p.cell_cls.data(), p.cell_cls.size(), s);
        upload(c->row_cls, p.cls.data(), p.cls.size(), s);
        upload(c->tile_ptr, p.tile_ptr.data(), p.tile_ptr.size(), s);
        upload(c->tile_cell, p.tile_cell.data(), p.tile_cell.size(), s);
        upload(c->tile_cols, p.tile_cols.data(), p.tile_cols.size(), s);
        upload(c->tile_start, p.tile_start.data(), p.tile_start.size(), s);
        // cluster plan of the compressed inner operator (FP32-IR contexts)
        c->hmp = HmatPlan();
        if (c->cfg.precision == WB_PREC_FP32_IR) {
            plan_hmat(p, Xi.data(), c->hmp);
            upload(c->hm_idx, c->hmp.cl_idx.data(), c->hmp.cl_idx.size(), s);
            upload(c->hm_size, c->hmp.cl_size.data(), c->hmp.cl_size.size(), s);
            upload(c->hm_near, c->hmp.near.data(), c->hmp.near.size(), s);
            // row -> its position J HM_CL + p in the cluster order
            std::vector<int> pos(std::max(c->n, 1), -1);  // -1: in no cluster
            for (int J = 0; J < c->hmp.ncl; ++J)
                for (int q = 0; q < c->hmp.cl_size[J]; ++q)
                    pos[c->hmp.cl_idx[(size_t)J * HM_CL + q]] = J * HM_CL + q;
            upload(c->hm_pos, pos.data(), pos.size(), s);
        }
        c->hm_xc_for = nullptr;
        c->hm_assembly_ver = c->hm_mask_ver = -1;
        c->hm_agg = 0;
        if (p.n_tasks > 0) {
            upload(c->st_task, p.st_task.data(), p.st_task.size(), s);
            upload(c->st_gap, p.st_gap.data(), p.st_gap.size(), s);
            upload(c->st_gcell, p.st_gcell.data(), p.st_gcell.size(), s);
        }
        upload(c->cell_int_s, p.cell_int_s.data(), p.cell_int_s.size(), s);
        {
            std::vector<int8_t> os(nc);
            for (int k = 0; k < nc; ++k) os[k] = (int8_t)(cell_orient[k] * p.canon_sign[k]);
            upload(c->orient_s, os.data(), os.size(), s);
        }
        upload(c->sing_ptr, p.sing_ptr.data(), p.sing_ptr.size(), s);
        upload(c->sing_cell, p.sing_cell.data(), p.sing_cell.size(), s);
        upload(c->sing_vtx, p.sing_vtx.data(), p.sing_vtx.size(), s);
        upload(c->perm_d, p.perm.data(), p.perm.size(), s);
        upload(c->iperm_d, p.iperm.data(), p.iperm.size(), s);
        {
            // node -> (cell, local vertex) incidence in cell order, for the L2 projections
            std::vector<int> ip(n + 1, 0), it(4 * (size_t)nc);
            for (size_t k = 0; k < 4 * (size_t)nc; ++k) ip[p.cell_int[k] + 1]++;
            for (int i = 0; i < n; ++i) ip[i + 1] += ip[i];
            std::vector<int> fill(ip.begin(), ip.end() - 1);
            for (int cc = 0; cc < nc; ++cc)
                for (int k = 0; k < 4; ++k) it[fill[p.cell_int[4 * cc + k]]++] = 4 * cc + k;
            upload(c->inc_ptr, ip.data(), ip.size(), s);
            upload(c->inc_item, it.data(), it.size(), s);
        }
        const Quadrature &q = c->quad;
        upload(c->qpts, q.reg_pts.data(), q.reg_pts.size(), s);
        upload(c->qw, q.reg_w.data(), q.reg_w.size(), s);
        const int ns = 2 * q.s_order * q.s_order;
        std::vector<double> rule(4 * (size_t)ns * 3);
        for (int v = 0; v < 4; ++v)
            for (int k = 0; k < ns; ++k) {
                rule[3 * (v * ns + k)] = q.sing_pts[2 * (v * ns + k)];
                rule[3 * (v * ns + k) + 1] = q.sing_pts[2 * (v * ns + k) + 1];
                rule[3 * (v * ns + k) + 2] = q.sing_w[v * ns + k];
            }
        upload(c->srule, rule.data(), rule.size(), s);
        const int nq = q.q_order * q.q_order;
        c->cellq.ensure(sizeof(double) * 8 * (size_t)nq * nc);
        c->cellq_s.ensure(sizeof(double) * 8 * (size_t)nq * (p.st_gcell.size() + 1));
        cell_records(c);
        alloc_matrices(c);
        c->hsum_part.ensure(sizeof(double) * (size_t)hsum_parts(c) * n);
        c->hsum_sing.ensure(sizeof(double) * n);
        c->alpha.ensure(sizeof(double) * c->vlen);
        HIPCHK(hipMemsetAsync(c->alpha.p, 0, c->alpha.bytes, s));
        c->surf_d.ensure(sizeof(double) * c->vlen);
        c->gmask.ensure((c->vlen + 3) / 4);
        // default: no constraints
        c->n_lines = 0;
        c->n_ct = 0;  // no transposed lines either (the old ones index the old mesh)
        std::vector<int> fr(n);
        for (int i = 0; i < n; ++i) fr[i] = i;
        c->n_free = n;
        upload(c->free_rows, fr.data(), fr.size(), s);
        std::vector<uint8_t> zf(n, 0);
        upload(c->constrained, zf.data(), zf.size(), s);
        std::vector<int> nol(std::max(n, 1), -1);
        upload(c->con_line, nol.data(), nol.size(), s);
        c->con_rows.ensure(4);
        c->con_ptr.ensure(4);
        c->con_cols.ensure(4);
        c->con_w.ensure(8);
        c->con_inhom.ensure(8);
        c->cons = ConstraintSet();
        c->cons.ptr.push_back(0);
        c->cons_auto = false;
        c->cons_struct_ver++;
        HIPCHK(hipStreamSynchronize(s));
        c->have_mesh = true;
    });
}

wb_status wb_set_support_points(wb_ctx *c, const double *support)
{
    return guarded(c, [&] {
        require(c->have_mesh, WB_E_STATE, "wb_set_support_points before wb_set_mesh");
        require(support != nullptr, WB_E_ARG, "null support points");
        const int n = c->n, nc = c->plan.nc;
        std::vector<double> Xi(3 * (size_t)n);
        for (int d = 0; d < n; ++d)
            for (int k = 0; k < 3; ++k) Xi[3 * c->plan.perm[d] + k] = support[3 * d + k];
        upload(c->X, Xi.data(), Xi.size(), c->stream);
        cell_records(c);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->assembled = false;
    });
}

wb_status wb_compute_normals(wb_ctx *c, int normalize, double *normals, int *cg_iterations)
{
    return guarded(c, [&] {
        require(c->have_mesh, WB_E_STATE, "wb_compute_normals before wb_set_mesh");
        require(normals != nullptr, WB_E_ARG, "null normals");
        project_fields(c, nullptr, normalize, normals, nullptr);
        if (cg_iterations) *cg_iterations = c->proj_iterations;
    });
}

wb_status wb_compute_surface_gradients(wb_ctx *c, const double *tmp_rhs, double *grads,
                                       int *cg_iterations)
{
    return guarded(c, [&] {
        require(c->have_mesh && c->have_masks, WB_E_STATE,
                "wb_compute_surface_gradients needs a mesh and masks");
        require(tmp_rhs && grads, WB_E_ARG, "null vector");
        std::vector<double> phi(c->n);
        for (int i = 0; i < c->n; ++i) phi[i] = tmp_rhs[i] * c->surface[i];
        project_fields(c, phi.data(), 0, nullptr, grads);
        if (cg_iterations) *cg_iterations = c->proj_iterations;
    });
}

wb_status wb_compute_potential_gradients(wb_ctx *c, const double *phi, const double *dphi_dn,
                                         double *grads, int *cg_iterations)
{
    return guarded(c, [&] {
        require(c->have_mesh, WB_E_STATE, "wb_compute_potential_gradients before wb_set_mesh");
        require(phi && dphi_dn && grads, WB_E_ARG, "null vector");
        project_fields(c, phi, 0, nullptr, grads, dphi_dn);
        // coincident dofs take the mean of their projected gradients
        const MeshPlan &p = c->plan;
        std::vector<double> g(grads, grads + 3 * (size_t)c->n);
        for (int i = 0; i < c->n; ++i) {
            const int q0 = p.dn_ptr[i], q1 = p.dn_ptr[i + 1];
            if (q1 - q0 < 2) continue;
            for (int e = 0; e < 3; ++e) {
                double a = 0.0;
                for (int q = q0; q < q1; ++q) a += g[3 * (size_t)p.dn_idx[q] + e];
                grads[3 * (size_t)i + e] = a / (q1 - q0);
            }
        }
        if (cg_iterations) *cg_iterations = c->proj_iterations;
    });
}

// G and H (and the lo halves in FP32_IR), zero-filled so the padding columns read as 0
static void alloc_matrices(wb_ctx *c)
{
    const size_t entries = (size_t)c->n * c->ld;
    c->G.ensure(entries * elem_bytes(c));
    c->H.ensure(entries * elem_bytes(c));
    HIPCHK(hipMemsetAsync(c->G.p, 0, c->G.bytes, c->stream));
    HIPCHK(hipMemsetAsync(c->H.p, 0, c->H.bytes, c->stream));
    if (c->cfg.precision == WB_PREC_FP32_IR) {
        c->Glo.ensure(entries * 2);
        c->Hlo.ensure(entries * 2);
        HIPCHK(hipMemsetAsync(c->Glo.p, 0, c->Glo.bytes, c->stream));
        HIPCHK(hipMemsetAsync(c->Hlo.p, 0, c->Hlo.bytes, c->stream));
    } else {
        c->Glo.release();
        c->Hlo.release();
    }
}

// comp_dom.surface_nodes / other_nodes: host copies, device mask, column groups
static void install_masks(wb_ctx *c, const double *surface, const double *other)
{
    const int n = c->n;
    // vmult/compute_rhs use the masks as multipliers; they must be complementary 0/1
    for (int i = 0; i < n; ++i)
        require((surface[i] == 0.0 || surface[i] == 1.0) && other[i] == 1.0 - surface[i],
                WB_E_UNSUPPORTED, "masks must be complementary 0/1 vectors");
    c->surface.assign(surface, surface + n);
    c->other.assign(other, other + n);
    double linf = 0;
    for (int i = 0; i < n; ++i) linf = std::max(linf, std::fabs(surface[i]));
    c->pure_neumann = linf < 1e-10;
    std::vector<double> si;
    to_internal(c, surface, si);
    upload(c->surf_d, si.data(), c->vlen, c->stream);
    {
        std::vector<uint8_t> sc(std::max(n, 1), 0);  // caller order, for the solution scatter
        for (int i = 0; i < n; ++i) sc[i] = surface[i] != 0.0;
        upload(c->surf_c, sc.data(), sc.size(), c->stream);
    }
    // column selection: bit set -> column comes from H in vmult (other_nodes != 0)
    std::vector<double> oi;
    to_internal(c, other, oi);
    const size_t ng = (c->vlen + 3) / 4;
    std::vector<uint8_t> gm(ng, 0);
    for (int j = 0; j < n; ++j)
        if (oi[j] != 0.0) gm[j / 4] |= (uint8_t)(1u << (j % 4));
    upload(c->gmask, gm.data(), gm.size(), c->stream);
    // runs of 4-column groups reading only G, only H, or mixed
    std::vector<int> sg;
    const int ngroups = (n + 3) / 4;
    auto kind = [&](int gg) {
        const int nv = std::min(4, n - 4 * gg);
        const uint8_t f = (uint8_t)((1u << nv) - 1);
        const uint8_t m = gm[gg] & f;
        return m == 0 ? 0 : (m == f ? 1 : 2);
    };
    for (int g = 0; g < ngroups;) {
        const int k = kind(g);
        int e = g + 1;
        while (e < ngroups && kind(e) == k) ++e;
        sg.insert(sg.end(), {g, e, k, 0});
        g = e;
    }
    c->n_segs = (int)sg.size() / 4;
    upload(c->segs, sg.data(), sg.size(), c->stream);
    // segments of the compressed operator's clusters: {first column, length, 0 = all
    // from G, 1 = all from H, 2 = mixed, 0}
    if (c->hmp.ncl > 0) {
        std::vector<int> hs((size_t)c->hmp.ncl * HM_SEGS * 4, 0);
        for (size_t q = 0; q < (size_t)c->hmp.ncl * HM_SEGS; ++q) {
            const int a = c->hmp.cl_seg[2 * q], len = c->hmp.cl_seg[2 * q + 1];
            int nh = 0;
            for (int j = a; j < a + len; ++j) nh += oi[j] != 0.0;
            hs[4 * q] = a;
            hs[4 * q + 1] = len;
            hs[4 * q + 2] = nh == 0 ? 0 : (nh == len ? 1 : 2);
        }
        upload(c->hm_seg, hs.data(), hs.size(), c->stream);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_masks = true;
    c->mask_ver++;
    if (c->assembled) refresh_diag(c);
}

wb_status wb_prepare_bem_vectors(wb_ctx *c, int n_fs_ids, const int32_t *fs_ids,
                                 int n_wall_ids, const int32_t *wall_ids, int n_inflow_ids,
                                 const int32_t *inflow_ids, const double *wind,
                                 const double *inflow_dphi_dn, const uint8_t *transom,
                                 const double *prev_dphi_dn, double *bem_bc, double *dphi_dn)
{
    return guarded(c, [&] {
        require(c->have_mesh, WB_E_STATE, "wb_prepare_bem_vectors before wb_set_mesh");
        require(!c->cell_mat_h.empty(), WB_E_STATE,
                "wb_prepare_bem_vectors needs the cell materials of wb_set_mesh");
        require(bem_bc && dphi_dn && wind, WB_E_ARG, "null vector");
        require(n_fs_ids >= 0 && n_wall_ids >= 0 && n_inflow_ids >= 0, WB_E_ARG,
                "negative id count");
        require((n_fs_ids == 0 || fs_ids) && (n_wall_ids == 0 || wall_ids) &&
                    (n_inflow_ids == 0 || inflow_ids),
                WB_E_ARG, "null id list");
        require(!transom || prev_dphi_dn, WB_E_ARG, "transom flags need prev_dphi_dn");
        auto in = [](int m, int k, const int32_t *ids) {
            for (int q = 0; q < k; ++q)
                if (ids[q] == m) return true;
            return false;
        };
        const int n = c->n, nc = c->plan.nc;
        const int *cd = c->cell_dofs_h.data();
        const int *mat = c->cell_mat_h.data();
        // masks: 1 on the dofs of free-surface cells (free_surface.cc:8753-8781)
        std::vector<double> surf(n, 0.0), oth(n, 1.0);
        for (int k = 0; k < nc; ++k)
            if (in(mat[k], n_fs_ids, fs_ids))
                for (int v = 0; v < 4; ++v) {
                    surf[cd[4 * k + v]] = 1.0;
                    oth[cd[4 * k + v]] = 0.0;
                }
        // comp_dom.node_normals: the unnormalised projection
        // (NumericalTowingTank::compute_normals_at_nodes) at the current support points
        std::vector<double> nn(3 * (size_t)n);
        project_fields(c, nullptr, 0, nn.data(), nullptr);
        // boundary values on the dofs of the other cells, in cell order (:8787-8847)
        for (int k = 0; k < nc; ++k) {
            if (in(mat[k], n_fs_ids, fs_ids)) continue;
            for (int v = 0; v < 4; ++v) {
                const int j = cd[4 * k + v];
                double b = 0.0;
                if (in(mat[k], n_wall_ids, wall_ids)) {
                    const double *w = wind + 3 * (size_t)j, *q = nn.data() + 3 * (size_t)j;
                    b = -(q[0] * w[0] + q[1] * w[1] + q[2] * w[2]);
                } else if (in(mat[k], n_inflow_ids, inflow_ids)) {
                    b = inflow_dphi_dn ? inflow_dphi_dn[j] : 0.0;
                }
                bem_bc[j] = b;
                dphi_dn[j] = b;
            }
        }
        // transom stern nodes on the water: Neumann, with the mean of the previous dphi_dn
        // of their coincident dofs (:8856-8880)
        if (transom)
            for (int i = 0; i < n; ++i) {
                if (!transom[i]) continue;
                surf[i] = 0.0;
                oth[i] = 1.0;
                const int p0 = c->plan.dn_ptr[i], p1 = c->plan.dn_ptr[i + 1];
                int cnt = 0;
                for (int p = p0; p < p1; ++p) cnt += c->plan.dn_idx[p] != i;
                double b = 0.0;
                for (int p = p0; p < p1; ++p)
                    if (c->plan.dn_idx[p] != i) b += prev_dphi_dn[c->plan.dn_idx[p]] / cnt;
                bem_bc[i] = b;
                dphi_dn[i] = b;
            }
        install_masks(c, surf.data(), oth.data());
    });
}

wb_status wb_get_masks(wb_ctx *c, double *surface, double *other)
{
    return guarded(c, [&] {
        require(c->have_masks, WB_E_STATE, "no masks installed");
        if (surface) std::copy(c->surface.begin(), c->surface.end(), surface);
        if (other) std::copy(c->other.begin(), c->other.end(), other);
    });
}

wb_status wb_set_masks(wb_ctx *c, const double *surface, const double *other)
{
    return guarded(c, [&] {
        require(c->have_mesh, WB_E_STATE, "wb_set_masks before wb_set_mesh");
        require(surface && other, WB_E_ARG, "null mask");
        install_masks(c, surface, other);
    });
}

wb_status wb_assemble(wb_ctx *c)
{
    return guarded(c, [&] {
        require(c->have_mesh, WB_E_STATE, "wb_assemble before wb_set_mesh");
        HIPCHK(hipSetDevice(c->cfg.device));
        if (!c->G.p) alloc_matrices(c);  // after wb_release_matrices
        const MeshPlan &p = c->plan;
        PhiTable nodes{};
        for (int i = 0; i < c->quad.q_order; ++i) nodes.phi[i] = c->quad.reg_pts[2 * i];
        {
            KernelScope ks(c, WB_K_ASSEMBLE, c->n);
            launch_assemble_strips(c->cfg.precision, c->quad.q_order, c->X.as<double>(),
                                   c->cellq_s.as<double>(), c->st_task.as<int>(),
                                   c->st_gap.as<int>(), c->n, p.n_tasks, c->ld, nodes, c->G.p,
                                   c->H.p, c->Glo.p, c->Hlo.p, c->hsum_part.as<double>(),
                                   c->stream);
            HIPCHK(hipGetLastError());
            if (p.n_tiles > 0)
                launch_assemble_tiles(c->cfg.precision, c->quad.q_order, p.tc, c->X.as<double>(),
                                      c->cellq.as<double>(), c->tile_ptr.as<int>(),
                                      c->tile_cell.as<int>(), c->tile_cols.as<int>(),
                                      c->tile_start.as<int>(), c->n, p.n_tiles, c->ld, nodes,
                                      c->G.p, c->H.p, c->Glo.p, c->Hlo.p,
                                      c->hsum_part.as<double>() + (size_t)p.n_tasks * c->n,
                                      c->stream);
            HIPCHK(hipGetLastError());
        }
        const int ns = 2 * c->quad.s_order * c->quad.s_order;
        {
            KernelScope ks(c, WB_K_SINGULAR, c->n);
            launch_assemble_singular(c->cfg.precision, c->X.as<double>(), c->cell_int.as<int>(),
                                     c->orient.as<int8_t>(), c->sing_ptr.as<int>(),
                                     c->sing_cell.as<int>(), c->sing_vtx.as<int>(), c->n, ns,
                                     c->srule.as<double>(), c->cellq.as<double>(),
                                     c->quad.q_order * c->quad.q_order, c->qpts.as<double>(),
                                     c->ld, c->G.p, c->H.p, c->Glo.p, c->Hlo.p,
                                     c->hsum_sing.as<double>(), c->stream);
            HIPCHK(hipGetLastError());
        }
        launch_alpha_reduce(c->hsum_part.as<double>(), hsum_parts(c), c->hsum_sing.as<double>(),
                            c->n, c->alpha.as<double>(), c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        c->assembled = true;
        c->assembly_ver++;
        if (c->have_masks) refresh_diag(c);
    });
}

wb_status wb_release_matrices(wb_ctx *c)
{
    return guarded(c, [&] {
        require(c->have_mesh, WB_E_STATE, "wb_release_matrices before wb_set_mesh");
        HIPCHK(hipSetDevice(c->cfg.device));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (DevBuf *b : {&c->G, &c->H, &c->Glo, &c->Hlo, &c->A16, &c->cs_S}) b->release();
        c->assembled = false;
    });
}

// rows[r] (caller numbering) of G and H, caller column order, FP64 (hi + lo in FP32_IR)
static void copy_rows(wb_ctx *c, int nr, const int *rows, double *G, double *H)
{
    require(c->assembled, WB_E_STATE, "matrix copy before wb_assemble");
    HIPCHK(hipStreamSynchronize(c->stream));
    const int n = c->n;
    const size_t ld = c->ld;
    for (int which = 0; which < 2; ++which) {
        double *out = which == 0 ? G : H;
        if (!out) continue;
        const DevBuf &hi = which == 0 ? c->G : c->H;
        const DevBuf &lo = which == 0 ? c->Glo : c->Hlo;
        std::vector<double> row(ld);
        std::vector<float> rowf(ld);
        std::vector<uint16_t> rowl(ld);
        for (int r = 0; r < nr; ++r) {
            const int d = rows ? rows[r] : r;
            require(d >= 0 && d < n, WB_E_ARG, "row index out of range");
            const int i = c->plan.perm[d];
            if (c->cfg.precision == WB_PREC_FP64) {
                HIPCHK(hipMemcpy(row.data(), (char *)hi.p + sizeof(double) * ld * i,
                                 sizeof(double) * ld, hipMemcpyDeviceToHost));
            } else {
                HIPCHK(hipMemcpy(rowf.data(), (char *)hi.p + sizeof(float) * ld * i,
                                 sizeof(float) * ld, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(rowl.data(), (char *)lo.p + sizeof(uint16_t) * ld * i,
                                 sizeof(uint16_t) * ld, hipMemcpyDeviceToHost));
                for (size_t e = 0; e < ld; ++e) {
                    uint32_t b = (uint32_t)rowl[e] << 16;
                    float lf;
                    std::memcpy(&lf, &b, 4);
                    row[e] = (double)rowf[e] + (double)lf;
                }
            }
            double *o = out + (size_t)r * n;
            for (int e = 0; e < n; ++e) o[e] = row[c->plan.perm[e]];
        }
    }
}

wb_status wb_get_matrices(wb_ctx *c, double *G, double *H)
{
    return guarded(c, [&] { copy_rows(c, c->n, nullptr, G, H); });
}

wb_status wb_get_matrix_rows(wb_ctx *c, int n_rows, const int32_t *rows, double *G, double *H)
{
    return guarded(c, [&] {
        require(n_rows >= 0 && (rows || n_rows == 0), WB_E_ARG, "bad row list");
        copy_rows(c, n_rows, rows, G, H);
    });
}

wb_status wb_get_coarse_matrix(wb_ctx *c, int *nc, int *agg, double *Ac)
{
    return guarded(c, [&] {
        require(nc && agg, WB_E_ARG, "null output");
        require(c->cs_on, WB_E_STATE, "no coarse space in use (build one with a solve)");
        *nc = c->cs_nc;
        *agg = c->cs_agg;
        if (Ac) {
            HIPCHK(hipMemcpyAsync(Ac, c->cs_Ac.p, sizeof(double) * c->cs_nc * c->cs_nc,
                                  hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
    });
}

wb_status wb_compute_alpha(wb_ctx *c, double *alpha)
{
    return guarded(c, [&] {
        require(c->assembled, WB_E_STATE, "wb_compute_alpha before wb_assemble");
        require(alpha != nullptr, WB_E_ARG, "null alpha");
        download_vec(c, c->alpha, alpha);
    });
}

static void vmult_like(wb_ctx *c, int which, double *dst, const double *src)
{
    check_ready(c, true);
    require(dst && src, WB_E_ARG, "null vector");
    ensure_work(c);
    upload_vec(c, c->vs, src);
    const int mode = c->cfg.precision == WB_PREC_FP64 ? 0 : 2;
    gemv(c, mode, which, c->vs.p, c->vs.p, c->vr.p, false);
    if (which == 0 && c->pure_neumann) {
        const double nrm = device_norm(c, c->vr.p, 0);
        std::vector<double> v(c->vlen);
        HIPCHK(hipMemcpy(v.data(), c->vr.p, sizeof(double) * c->vlen, hipMemcpyDeviceToHost));
        for (int i = 0; i < c->n; ++i) v[i] -= nrm;
        to_caller(c, v, dst);
        return;
    }
    download_vec(c, c->vr, dst);
}

wb_status wb_vmult(wb_ctx *c, double *dst, const double *src)
{
    return guarded(c, [&] { vmult_like(c, 0, dst, src); });
}

wb_status wb_compute_rhs(wb_ctx *c, double *dst, const double *src)
{
    return guarded(c, [&] { vmult_like(c, 1, dst, src); });
}

wb_status wb_set_constraints(wb_ctx *c, int n_lines, const int32_t *rows, const int32_t *ptr,
                             const int32_t *cols, const double *w, const double *inhom)
{
    return guarded(c, [&] {
        require(c->have_mesh, WB_E_STATE, "wb_set_constraints before wb_set_mesh");
        require(n_lines >= 0, WB_E_ARG, "negative n_lines");
        ConstraintSet cs;
        cs.ptr.push_back(0);
        if (n_lines > 0) {
            require(rows && ptr && inhom, WB_E_ARG, "null constraint arrays");
            require(ptr[0] == 0, WB_E_ARG, "line_ptr[0] must be 0");
            const int nnz = ptr[n_lines];
            require(nnz >= 0 && (nnz == 0 || (cols && w)), WB_E_ARG, "bad constraint entries");
            cs.rows.assign(rows, rows + n_lines);
            cs.ptr.assign(ptr, ptr + n_lines + 1);
            for (int l = 0; l < n_lines; ++l)
                require(cs.ptr[l + 1] >= cs.ptr[l], WB_E_ARG, "line_ptr decreasing");
            cs.cols.assign(cols, cols + nnz);
            cs.w.assign(w, w + nnz);
            cs.inhom.assign(inhom, inhom + n_lines);
        }
        c->cons_auto = false;
        install_constraints(c, cs);
    });
}

wb_status wb_compute_constraints(wb_ctx *c, const double *tmp_rhs, const double *normals,
                                 const double *grads, int *n_lines)
{
    return guarded(c, [&] {
        check_ready(c, false);
        require(tmp_rhs != nullptr, WB_E_ARG, "null tmp_rhs");
        c->cons_normals.clear();
        c->cons_grads.clear();
        if (normals) c->cons_normals.assign(normals, normals + 3 * (size_t)c->n);
        if (grads) c->cons_grads.assign(grads, grads + 3 * (size_t)c->n);
        c->cons_project = !normals || !grads;
        c->cons_auto = true;
        refresh_auto_constraints(c, tmp_rhs);
        if (n_lines) *n_lines = c->n_lines;
    });
}

wb_status wb_constrained_vmult(wb_ctx *c, double *dst, const double *src)
{
    return guarded(c, [&] {
        check_ready(c, true);
        require(dst && src, WB_E_ARG, "null vector");
        ensure_work(c);
        upload_vec(c, c->vs, src);
        apply_constrained(c, c->cfg.precision == WB_PREC_FP64 ? 0 : 2, c->vs.p, c->vr.p);
        download_vec(c, c->vr, dst);
    });
}

wb_status wb_inner_operator_info(wb_ctx *c, int64_t *info)
{
    return guarded(c, [&] {
        check_ready(c, true);
        require(info != nullptr, WB_E_ARG, "null output");
        require(c->cfg.precision == WB_PREC_FP32_IR && c->cfg.ir_inner_format == 2,
                WB_E_UNSUPPORTED, "FP32_IR contexts with ir_inner_format 2 only");
        ensure_hm(c);
        const int ncl = c->hmp.ncl;
        std::vector<uint8_t> rk((size_t)ncl * ncl);
        HIPCHK(hipMemcpyAsync(rk.data(), c->hm_rank.p, rk.size(), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        int64_t nd = 0, nl = 0, nz = 0, rs = 0, bytes = 0;
        for (int I = 0; I < ncl; ++I)
            for (int J = 0; J < ncl; ++J) {
                const int r = rk[(size_t)I * ncl + J];
                const int64_t a = c->hmp.cl_size[I], b = c->hmp.cl_size[J];
                if (r == HM_DENSE) {
                    ++nd;
                    bytes += 2 * a * b;
                } else if (r == 0) {
                    ++nz;
                } else {
                    ++nl;
                    rs += r;
                    bytes += 2 * (int64_t)r * (a + b);
                }
            }
        const int64_t v[6] = {ncl, nd, nl, nz, rs, bytes};
        std::copy(v, v + 6, info);
    });
}

wb_status wb_inner_vmult(wb_ctx *c, double *dst, const double *src)
{
    return guarded(c, [&] {
        check_ready(c, true);
        require(dst && src, WB_E_ARG, "null vector");
        require(c->cfg.precision == WB_PREC_FP32_IR, WB_E_UNSUPPORTED,
                "the inner operator exists in FP32_IR contexts only");
        ensure_work(c);
        const int fmode = c->cfg.ir_inner_format == 2 ? 4 : (c->cfg.ir_inner_format == 1 ? 3 : 1);
        if (fmode == 3) ensure_h16(c);
        if (fmode == 4) ensure_hm(c);
        const size_t vb4 = sizeof(float) * c->vlen;
        for (DevBuf *b : {&c->z32, &c->w32}) {
            if (b->bytes < vb4) {
                b->ensure(vb4);
                HIPCHK(hipMemsetAsync(b->p, 0, vb4, c->stream));
            }
        }
        upload_vec(c, c->vs, src);
        launch_vec_ops(1, c->n, c->vs.p, nullptr, c->z32.p, 0, c->stream);
        apply_constrained(c, fmode, c->z32.p, c->w32.p);
        launch_vec_ops(6, c->n, nullptr, nullptr, c->vr.p, 0, c->stream);
        launch_vec_ops(2, c->n, c->w32.p, nullptr, c->vr.p, 0, c->stream);
        download_vec(c, c->vr, dst);
    });
}

wb_status wb_constrained_vmult_noalloc(wb_ctx *c, const double *src, double *dst)
{
    return guarded(c, [&] {
        check_ready(c, true);
        require(src && dst, WB_E_ARG, "null vector");
        require(c->vlen <= (size_t)NOALLOC_MAX_N, WB_E_ARG,
                "wb_constrained_vmult_noalloc: more than 262144 dofs (rounded up to 64) "
                "exceed its fixed work space");
        require(!c->pure_neumann, WB_E_UNSUPPORTED,
                "wb_constrained_vmult_noalloc: pure-Neumann operators need a host norm");
        require(!c->op_transpose, WB_E_STATE, "transposed operator active");
        double *ws = noalloc_scratch(c->cfg.device);
        require(ws != nullptr, WB_E_STATE, "no work space on this device");
        double *xs = ws, *ys = ws + NOALLOC_MAX_N;  // source and product, device order
        hipStream_t s = c->stream;
        // the array keeps the last call's contents: clear the source's padding columns
        launch_vec_ops(6, (int)c->vlen, nullptr, nullptr, xs, 0, s);
        launch_permute(c->n, c->iperm_d.as<int>(), src, xs, 1, s);
        const int mode = c->cfg.precision == WB_PREC_FP64 ? 0 : 2;
        launch_gemv_sel(mode, c->G.p, c->H.p, c->Glo.p, c->Hlo.p, c->ld, c->segs.as<int>(),
                        c->n_segs, c->gmask.as<uint8_t>(), 0, 1.0, -1.0, xs,
                        c->diag_op.as<double>(), xs, c->free_rows.as<int>(), c->n_free, ys, s);
        launch_constraint_rows(0, c->n_lines, c->con_rows.as<int>(), c->con_ptr.as<int>(),
                               c->con_cols.as<int>(), c->con_w.as<double>(),
                               c->con_inhom.as<double>(), 0, xs, ys, s);
        launch_permute(c->n, c->perm_d.as<int>(), ys, dst, 1, s);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    });
}

wb_status wb_solve_system(wb_ctx *c, const double *tmp_rhs, double *phi, double *dphi_dn,
                          wb_solve_stats *st)
{
    return guarded(c, [&] {
        check_ready(c, true);
        require(tmp_rhs && phi && dphi_dn, WB_E_ARG, "null vector");
        refresh_auto_constraints(c, tmp_rhs);
        ensure_work(c);
        upload_vec(c, c->vs, tmp_rhs);
        solve_system_dev(c, c->vs.as<double>(), st);
        std::vector<double> sol(c->n);
        download_vec(c, c->vx, sol.data());
        for (int i = 0; i < c->n; ++i) {
            if (c->surface[i] == 0)
                phi[i] = sol[i];
            else
                dphi_dn[i] = sol[i];
        }
    });
}

wb_status wb_solve_system_dev(wb_ctx *c, const double *tmp_dev, double *phi_dev,
                              double *dphi_dev, wb_solve_stats *st)
{
    return guarded(c, [&] {
        check_ready(c, true);
        require(tmp_dev && phi_dev && dphi_dev, WB_E_ARG, "null vector");
        ensure_work(c);
        // caller order -> internal order on the device
        launch_permute(c->n, c->iperm_d.as<int>(), tmp_dev, c->vs.as<double>(), 1, c->stream);
        if (c->cons_auto) {
            std::vector<double> h(c->n);
            HIPCHK(hipMemcpyAsync(h.data(), tmp_dev, sizeof(double) * c->n,
                                  hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            refresh_auto_constraints(c, h.data());
        }
        solve_system_dev(c, c->vs.as<double>(), st);
        // scatter by mask (source/bem_problem.cc:869-878): internal sol -> caller phi where
        // the surface mask is 0, dphi_dn where it is 1
        launch_permute(c->n, c->perm_d.as<int>(), c->vx.as<double>(), c->vt.as<double>(), 1,
                       c->stream);  // vt[d] = sol[perm[d]] (caller order)
        launch_scatter_mask(c->n, c->surf_c.as<uint8_t>(), c->vt.as<double>(), phi_dev, dphi_dev,
                            c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

wb_status wb_solve(wb_ctx *c, const double *tmp_rhs, double *phi, double *dphi_dn,
                   wb_solve_stats *st)
{
    wb_status s = wb_assemble(c);
    if (s != WB_OK) return s;
    return wb_solve_system(c, tmp_rhs, phi, dphi_dn, st);
}

wb_status wb_residual(wb_ctx *c, double *res, const double *phi, const double *dphi_dn)
{
    return guarded(c, [&] {
        check_ready(c, true);
        require(res && phi && dphi_dn, WB_E_ARG, "null vector");
        const int n = c->n;
        std::vector<double> tmp(n), sol(n);
        for (int i = 0; i < n; ++i) {
            tmp[i] = 0. + dphi_dn[i] * c->other[i] + phi[i] * c->surface[i];
            sol[i] = 0. + dphi_dn[i] * c->surface[i] + phi[i] * c->other[i];
        }
        refresh_auto_constraints(c, tmp.data());
        ensure_work(c);
        upload_vec(c, c->vs, tmp.data());
        build_rhs(c, c->vs.p);  // rrhs in vb
        upload_vec(c, c->vs, sol.data());
        apply_constrained(c, c->cfg.precision == WB_PREC_FP64 ? 0 : 2, c->vs.p, c->vr.p);
        launch_vec_ops(0, n, c->vr.p, c->vb.p, c->vt.p, 0, c->stream);  // res = A sol - rrhs
        download_vec(c, c->vt, res);
    });
}

// ------------------------------------------------------------------ adjoint
wb_status wb_solve_adjoint(wb_ctx *c, const double *rhs, double *lambda, wb_solve_stats *st)
{
    return guarded(c, [&] {
        check_ready(c, true);
        require(rhs && lambda, WB_E_ARG, "null adjoint vector");
        require(!c->pure_neumann, WB_E_UNSUPPORTED,
                "the adjoint solve does not support all-Neumann masks");
        ensure_work(c);
        upload_vec(c, c->vb, rhs);
        struct Flip {
            wb_ctx *c;
            Flip(wb_ctx *c_) : c(c_) { c->op_transpose = true; }
            ~Flip() { c->op_transpose = false; }
        };
        if (c->cfg.precision == WB_PREC_FP32_IR) {
            // iterative refinement of the transposed system: residuals with the transposed
            // hi + lo operator, inner GMRES on the transposed FP32 hi part, transposed block
            // Jacobi (the transposed coarse space is built from FP64 matrices only)
            ensure_precond(c);
            Flip flip(c);
            wb_solve_stats s{};
            run_solver(c, &s);
            if (st) *st = s;
            download_vec(c, c->vx, lambda);
            return;
        }
        double t0 = now_ms();
        ensure_precond(c);
        if (c->cs_on && !c->cs_t_on) {
            // the coarse space of the transposed system, from the same aggregates and Ac
            c->cs_PTA.ensure(sizeof(double) * (size_t)c->cs_nc * c->vlen);
            c->cs_WTT.ensure(sizeof(double) * (size_t)c->cs_nc * c->vlen);
            launch_coarse_build_t(c->n, c->cs_agg, c->cs_nc, c->G.as<double>(), c->H.as<double>(),
                                  c->ld, c->alpha.as<double>(), c->gmask.as<uint8_t>(),
                                  c->constrained.as<uint8_t>(), c->cs_Aci.as<double>(),
                                  c->cs_PTA.as<double>(), c->vlen, c->cs_WTT.as<double>(),
                                  c->stream);
            HIPCHK(hipGetLastError());
            c->cs_t_on = true;
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        double t1 = now_ms();
        launch_vec_ops(6, c->n, nullptr, nullptr, c->vx.p, 0, c->stream);
        Flip flip(c);
        GmresResult R = gmres(c, 0, c->vb.p, c->vx.p, c->cfg.solver_tolerance, 0.0,
                              c->cfg.solver_max_steps, true);
        HIPCHK(hipStreamSynchronize(c->stream));
        double t2 = now_ms();
        wb_solve_stats s{};
        s.iterations = R.iterations;
        s.restarts = R.restarts;
        s.residual = R.residual;
        s.t_precond_ms = t1 - t0;
        s.t_gmres_ms = t2 - t1;
        // true relative residual of the transposed system
        apply_constrained(c, 0, c->vx.p, c->w.p);
        launch_vec_ops(8, c->n, c->w.p, c->vb.p, c->vt.p, 0, c->stream);
        const double rn = device_norm(c, c->vt.p, 0), bn = device_norm(c, c->vb.p, 0);
        s.true_rel_residual = bn > 0 ? rn / bn : rn;
        if (st) *st = s;
        require(R.status == 1, WB_E_NOCONV, "adjoint GMRES did not converge");
        download_vec(c, c->vx, lambda);
    });
}

wb_status wb_residual_vjp(wb_ctx *c, const double *w, const double *phi, const double *dphi_dn,
                          int free_rows_only, double *grad)
{
    return guarded(c, [&] {
        require(c->assembled, WB_E_STATE, "wb_residual_vjp before wb_assemble");
        require(w && phi && dphi_dn && grad, WB_E_ARG, "null vjp argument");
        const int n = c->n, nc = c->plan.nc;
        const int nq = c->quad.q_order * c->quad.q_order;
        const int ns = 2 * c->quad.s_order * c->quad.s_order;
        require(nq <= 64, WB_E_UNSUPPORTED, "regular rule too large for the backward pass");
        std::vector<double> wi, pi, si;
        to_internal(c, w, wi);
        to_internal(c, phi, pi);
        to_internal(c, dphi_dn, si);
        if (free_rows_only)
            for (int l = 0; l < (int)c->cons.rows.size(); ++l) wi[c->plan.perm[c->cons.rows[l]]] = 0.0;
        // rowpack[i] = {x, y, z, w_i, phi_i, class}
        std::vector<double> xs_host(3 * (size_t)n);
        HIPCHK(hipMemcpyAsync(xs_host.data(), c->X.p, sizeof(double) * 3 * n,
                              hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<double> rp(8 * (size_t)n, 0.0);
        for (int i = 0; i < n; ++i) {
            for (int d = 0; d < 3; ++d) rp[8 * i + d] = xs_host[3 * i + d];
            rp[8 * i + 3] = wi[i];
            rp[8 * i + 4] = pi[i];
            rp[8 * i + 5] = (double)c->plan.cls[i];
        }
        upload(c->vjp_w, wi.data(), n, c->stream);
        upload(c->vjp_phi, pi.data(), n, c->stream);
        upload(c->vjp_psi, si.data(), n, c->stream);
        upload(c->vjp_rows, rp.data(), rp.size(), c->stream);
        const int npairs = (int)c->plan.sing_cell.size();
        c->vjp_cstate.ensure(sizeof(double) * 2 * (size_t)nc * nq);
        c->vjp_p1.ensure(sizeof(double) * 3 * (size_t)vjp_row_chunks(nc) * n);
        c->vjp_p2.ensure(sizeof(double) * 7 * (size_t)vjp_point_chunks(n) * nc * nq);
        c->vjp_xg.ensure(sizeof(double) * 3 * (size_t)n);
        c->vjp_vbar.ensure(sizeof(double) * 12 * (size_t)nc);
        c->vjp_xs.ensure(sizeof(double) * 3 * (size_t)n);
        c->vjp_spart.ensure(sizeof(double) * 12 * (size_t)std::max(npairs, 1));
        launch_residual_vjp(c->quad.q_order, n, nc, c->X.as<double>(), c->row_cls.as<int>(),
                            c->vjp_w.as<double>(), c->vjp_phi.as<double>(),
                            c->vjp_psi.as<double>(), c->vjp_rows.as<double>(),
                            c->cellq.as<double>(), c->cell_int.as<int>(), c->cell_cls.as<int>(),
                            c->orient.as<int8_t>(), c->qpts.as<double>(), c->qw.as<double>(),
                            c->sing_ptr.as<int>(), c->sing_cell.as<int>(), c->sing_vtx.as<int>(),
                            ns, c->srule.as<double>(), c->vjp_cstate.as<double>(),
                            c->vjp_p1.as<double>(), c->vjp_p2.as<double>(),
                            c->vjp_xg.as<double>(), c->vjp_vbar.as<double>(),
                            c->vjp_xs.as<double>(), c->vjp_spart.as<double>(), c->stream);
        HIPCHK(hipGetLastError());
        std::vector<double> xg(3 * (size_t)n), xsg(3 * (size_t)n), vbar(12 * (size_t)nc),
            sp(12 * (size_t)npairs);
        HIPCHK(hipMemcpyAsync(xg.data(), c->vjp_xg.p, sizeof(double) * xg.size(),
                              hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(xsg.data(), c->vjp_xs.p, sizeof(double) * xsg.size(),
                              hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(vbar.data(), c->vjp_vbar.p, sizeof(double) * vbar.size(),
                              hipMemcpyDeviceToHost, c->stream));
        if (npairs)
            HIPCHK(hipMemcpyAsync(sp.data(), c->vjp_spart.p, sizeof(double) * sp.size(),
                                  hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        // per support point (internal index), in a fixed order: row part, regular cell
        // vertices, singular cell vertices; then to the caller's numbering
        std::vector<double> g(3 * (size_t)n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) g[3 * i + d] = xg[3 * i + d] + xsg[3 * i + d];
        for (int cc = 0; cc < nc; ++cc)
            for (int k = 0; k < 4; ++k) {
                const int v = c->plan.cell_int[4 * cc + k];
                for (int d = 0; d < 3; ++d) g[3 * v + d] += vbar[12 * (size_t)cc + 3 * k + d];
            }
        for (int p = 0; p < npairs; ++p) {
            const int cc = c->plan.sing_cell[p];
            for (int k = 0; k < 4; ++k) {
                const int v = c->plan.cell_int[4 * cc + k];
                for (int d = 0; d < 3; ++d) g[3 * v + d] += sp[12 * (size_t)p + 3 * k + d];
            }
        }
        for (int dof = 0; dof < n; ++dof) {
            const int i = c->plan.perm[dof];
            for (int d = 0; d < 3; ++d) grad[3 * (size_t)dof + d] = g[3 * i + d];
        }
    });
}

wb_status wb_residual_jvp(wb_ctx *c, const double *dx, const double *phi, const double *dphi_dn,
                          int free_rows_only, double *dres)
{
    return guarded(c, [&] {
        require(c->assembled, WB_E_STATE, "wb_residual_jvp before wb_assemble");
        require(dx && phi && dphi_dn && dres, WB_E_ARG, "null jvp argument");
        const int n = c->n, nc = c->plan.nc;
        const int nq = c->quad.q_order * c->quad.q_order;
        const int ns = 2 * c->quad.s_order * c->quad.s_order;
        std::vector<double> pi, si, dxi(3 * (size_t)n);
        to_internal(c, phi, pi);
        to_internal(c, dphi_dn, si);
        for (int d = 0; d < n; ++d)
            for (int k = 0; k < 3; ++k) dxi[3 * c->plan.perm[d] + k] = dx[3 * (size_t)d + k];
        upload(c->vjp_phi, pi.data(), n, c->stream);
        upload(c->vjp_psi, si.data(), n, c->stream);
        upload(c->vjp_w, dxi.data(), dxi.size(), c->stream);  // the direction, 3 per point
        c->vjp_cstate.ensure(sizeof(double) * 2 * (size_t)nc * nq);
        c->vjp_p2.ensure(sizeof(double) * 8 * (size_t)nc * nq);              // tangent records
        c->vjp_p1.ensure(sizeof(double) * (size_t)vjp_row_chunks(nc) * n);   // row partials
        c->vjp_xs.ensure(sizeof(double) * n);                                // singular part
        c->vjp_xg.ensure(sizeof(double) * n);                                // result
        launch_residual_jvp(c->quad.q_order, n, nc, c->X.as<double>(), c->vjp_w.as<double>(),
                            c->row_cls.as<int>(), c->vjp_phi.as<double>(),
                            c->vjp_psi.as<double>(), c->cellq.as<double>(), c->cell_int.as<int>(),
                            c->cell_cls.as<int>(), c->orient.as<int8_t>(), c->qpts.as<double>(),
                            c->qw.as<double>(), c->sing_ptr.as<int>(), c->sing_cell.as<int>(),
                            c->sing_vtx.as<int>(), ns, c->srule.as<double>(),
                            free_rows_only && c->have_cons ? c->constrained.as<uint8_t>()
                                                           : nullptr,
                            c->vjp_cstate.as<double>(), c->vjp_p2.as<double>(),
                            c->vjp_p1.as<double>(), c->vjp_xs.as<double>(),
                            c->vjp_xg.as<double>(), c->stream);
        HIPCHK(hipGetLastError());
        std::vector<double> out(n);
        HIPCHK(hipMemcpyAsync(out.data(), c->vjp_xg.p, sizeof(double) * n, hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int d = 0; d < n; ++d) dres[d] = out[c->plan.perm[d]];
    });
}

wb_status wb_solve_rhs(wb_ctx *c, const double *rhs, double *sol, wb_solve_stats *st)
{
    return guarded(c, [&] {
        check_ready(c, true);
        require(rhs && sol, WB_E_ARG, "null vector");
        ensure_work(c);
        upload_vec(c, c->vb, rhs);
        wb_solve_stats s{};
        run_solver(c, &s);
        if (st) *st = s;
        download_vec(c, c->vx, sol);
    });
}

wb_status wb_solve_system_multi(wb_ctx *c, int k, const double *tmp_rhs, double *phi,
                                double *dphi_dn, const double *inhom_k, wb_solve_stats *st)
{
    return guarded(c, [&] {
        check_ready(c, true);
        require(k >= 1 && tmp_rhs && phi && dphi_dn, WB_E_ARG, "bad multi-rhs arguments");
        // batches of up to 16 right-hand sides share every matrix sweep
        wb_solve_stats acc{};
        const int nl = c->n_lines;
        for (int j0 = 0; j0 < k; j0 += WB_MULTI_MAX) {
            const int kb = std::min(WB_MULTI_MAX, k - j0);
            const size_t off = (size_t)j0 * c->n;
            wb_solve_stats s1{};
            solve_multi(c, kb, tmp_rhs + off, phi + off, dphi_dn + off,
                        inhom_k ? inhom_k + (size_t)j0 * nl : nullptr, &s1);
            acc.iterations += s1.iterations;
            acc.ir_sweeps = std::max(acc.ir_sweeps, s1.ir_sweeps);
            acc.residual = std::max(acc.residual, s1.residual);
            acc.true_rel_residual = std::max(acc.true_rel_residual, s1.true_rel_residual);
            acc.t_rhs_ms += s1.t_rhs_ms;
            acc.t_precond_ms += s1.t_precond_ms;
            acc.t_gmres_ms += s1.t_gmres_ms;
        }
        if (st) *st = acc;
    });
}

size_t wb_device_bytes(const wb_ctx *c)
{
    if (!c) return 0;
    size_t t = 0;
    for (const DevBuf *b : const_cast<wb_ctx *>(c)->all_bufs()) t += b->bytes;
    return t;
}

wb_status wb_synchronize(wb_ctx *c)
{
    return guarded(c, [&] { HIPCHK(hipStreamSynchronize(c->stream)); });
}

wb_status wb_set_profiling(wb_ctx *c, int enable)
{
    return guarded(c, [&] {
        fold_kernel_stats(c);
        c->kstats = wb_kernel_stats{};
        c->profiling = enable != 0;
        for (auto &k : c->kcalls) k = 0;
    });
}

wb_status wb_get_kernel_stats(wb_ctx *c, wb_kernel_stats *out)
{
    return guarded(c, [&] {
        require(out != nullptr, WB_E_ARG, "null stats");
        fold_kernel_stats(c);
        *out = c->kstats;
    });
}

}  // extern "C"
