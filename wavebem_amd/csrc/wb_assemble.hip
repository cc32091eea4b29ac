// wb_assemble.hip -- dense single-layer G and double-layer H assembly on gfx950.
//
// Replaces the loops of BEMProblem::assemble_system (source/bem_problem.cc:106-590):
//   G_ij = sum_{c ni j} sum_q  phi_{j,c}(q) JxW_cq / (4 pi |y_cq - x_i|)
//   H_ij = sum_{c ni j} sum_q  phi_{j,c}(q) (x_i - y_cq).n_cq JxW_cq / (4 pi |y_cq - x_i|^3)
// (LaplaceKernel::kernels, include/laplace_kernel.h:45-64), with the regular QGauss
// rule on cells that do not touch x_i and QGaussOneOverR on the cells that do
// (double-node membership test, source/bem_problem.cc:214-230).
//
// Regular cells: one wavefront owns 64 collocation rows (lane = row).  Cell quadrature
// data is wave-uniform and arrives through scalar loads; each lane forms the 4+4 local G/H
// values of the current cell for its row and adds them to register (G) and lane-private
// LDS (H) accumulators, which leave through LDS as whole row segments, written once.
//  * Strips (k_assemble_strips): on logically structured patches the wavefront sweeps a
//    run of grid lines of at most 8 dofs; every cell is evaluated once per strip task it
//    touches (about 1.1 times at 40k dofs) and its updates go to fixed slots.
//  * Tiles (k_assemble_tiles): elsewhere, column tiles of up to TILE_COLS dofs and every
//    cell touching the tile, with wave-uniform slot indices.
// Both evaluate every cell with the regular rule; the singular kernel replaces the
// regular part of each singular (row, cell) pair by its QGaussOneOverR integral.
#include "wb_internal.h"
#include "wb_device.h"

namespace wb {

__device__ __forceinline__ void split_hilo(double v, float &hi, uint16_t &lo)
{
    hi = (float)v;
    float r = (float)(v - (double)hi);
    // r is rounded to FP32 first, then to bf16 (nearest even, v_cvt_pk_bf16_f32 on
    // gfx950); the empty asm keeps the compiler from merging the two roundings into one
    // emulated FP64 -> bf16 rounding, which costs eight more instructions
    asm volatile("" : "+v"(r));
    lo = __builtin_bit_cast(uint16_t, (__bf16)r);
}

__device__ __forceinline__ double join_hilo(float hi, uint16_t lo)
{
    return (double)hi + (double)__uint_as_float((uint32_t)lo << 16);
}

// ------------------------------------------------------------------ cell geometry
// deal.II FEValues<2,3> with a Q1 mapping: x = sum phi_k v_k, a = dx/dxi, b = dx/deta,
// JxW = sqrt(det(J^T J)) w, n = orient (a x b)/|a x b|.
// Record: y(3), n*JxW(3), JxW, xi*JxW (xi = first reference coordinate of the point).
__global__ void k_cell_geometry(const double *__restrict__ X, const int *__restrict__ cell_int,
                                const int8_t *__restrict__ orient, int n_slots, int nq,
                                const double *__restrict__ qpts, const double *__restrict__ qw,
                                double *__restrict__ cellq, const int *__restrict__ slot_cell)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n_slots * nq) return;
    const int slot = gid / nq, q = gid % nq;
    const int c = slot_cell ? slot_cell[slot] : slot;
    if (c < 0) return;
    const double xi = qpts[2 * q], eta = qpts[2 * q + 1];
    const double phi[4] = {(1 - xi) * (1 - eta), xi * (1 - eta), (1 - xi) * eta, xi * eta};
    const double dxi[4] = {-(1 - eta), (1 - eta), -eta, eta};
    const double deta[4] = {-(1 - xi), -xi, (1 - xi), xi};
    double x[3] = {0, 0, 0}, a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        const int v = cell_int[4 * c + k];
        for (int d = 0; d < 3; ++d) {
            const double p = X[3 * v + d];
            x[d] += phi[k] * p;
            a[d] += dxi[k] * p;
            b[d] += deta[k] * p;
        }
    }
    const double g00 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    const double g01 = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    const double g11 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    const double jxw = sqrt(g00 * g11 - g01 * g01) * qw[q];
    double nn[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double nl = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    const double s = (orient[c] < 0 ? -1.0 : 1.0) / nl;
    double *r = cellq + (size_t)gid * 8;
    r[0] = x[0];
    r[1] = x[1];
    r[2] = x[2];
    r[3] = nn[0] * s * jxw;
    r[4] = nn[1] * s * jxw;
    r[5] = nn[2] * s * jxw;
    r[6] = jxw;
    r[7] = xi * jxw;
}

void launch_cell_geometry(const double *X, const int *cell_int, const int8_t *orient,
                          int n_slots, int nq, const double *qpts, const double *qw,
                          double *cellq, hipStream_t s, const int *slot_cell)
{
    const long total = (long)n_slots * nq;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_cell_geometry, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, X,
                       cell_int, orient, n_slots, nq, qpts, qw, cellq, slot_cell);
}

// ------------------------------------------------------------------ regular cells
// Q = points per direction of the tensor rule (NQ = Q*Q, point q = i + Q*j, i fastest).
// The Q1 shape values factor as phi_k(q) = a_k(xi_i) b_k(eta_j), so the four local sums
// are formed per rule row j from S0 = sum_i (1-xi_i) t_ij and S1 = sum_i xi_i t_ij:
// only the Q Gauss nodes are needed (kept in SGPRs).
//
// The TC tile columns of each lane's row: G in a TC-wide register vector, H in LDS.  A
// cell's four tile-local columns are wave-uniform, so G is updated with s_set_gpr_idx
// register moves and H with one no-return LDS add per column; both are lane-private, so
// there is no contention and the summation order is fixed.  Keeping H out of the registers
// is what lets four waves share a SIMD.
// Measured at 41k dofs (FP64 assembly): 3 waves per SIMD with both accumulator tiles in
// registers 67 ms; 4 waves with the H tile in LDS 51 ms (the few spilled registers of the
// 4-wave budget cost less than the lost occupancy).
#ifndef WB_ASM_WAVES
#define WB_ASM_WAVES 4  // requested waves per SIMD for the regular-cell kernel
#endif
#ifndef WB_ASM_JUNROLL
#define WB_ASM_JUNROLL 8  // unroll of the rule-row loop (>= Q: fully unrolled)
#endif
#ifndef WB_ASM_IUNROLL
#define WB_ASM_IUNROLL 8  // unroll of the points of one rule row
#endif
#ifndef WB_ASM_BATCH
#define WB_ASM_BATCH 4  // quadrature-point records per batch of scalar loads
#endif
#ifndef WB_ASM_HLDS
#define WB_ASM_HLDS 1  // H accumulators in LDS (1) or in registers like G (0)
#endif
#define WB_STR(x) #x
#define WB_UNROLL(n) _Pragma(WB_STR(unroll n))
// The four local G and H values of one cell for the lane's collocation point x:
// g_k = sum_q phi_k(q) JxW_q / |y_q - x|, h_k = sum_q phi_k(q) (y_q - x).N_q / |y_q - x|^3
// (N = n JxW; the 1/(4 pi) factors and the sign of H are applied at the write-out).
// rec: the cell's NQ records (wave-uniform, read by scalar loads).
// FAST: the values are stored as FP32 hi + bf16 lo, and 1/r takes one Newton step
// (rsqrt_f64_newton1) instead of the third-order one.
template <int Q, bool FAST = false>
__device__ __forceinline__ void cell_local_values(const double *rec, double x0, double x1,
                                                  double x2, const PhiTable &nodes,
                                                  double (&g)[4], double (&h)[4])
{
    double g0 = 0, g1 = 0, g2 = 0, g3 = 0, h0 = 0, h1 = 0, h2 = 0, h3 = 0;
    WB_UNROLL(WB_ASM_JUNROLL)
    for (int j = 0; j < Q; ++j) {
        double sg = 0, sg1 = 0, sh = 0, sh1 = 0;
        // The records of WB_ASM_BATCH points arrive in one batch of scalar loads:
        // scalar loads return out of order, so every use of one waits for all
        // outstanding ones (lgkmcnt(0)), and batching divides those waits.
        WB_UNROLL(WB_ASM_IUNROLL)
        for (int i = 0; i < Q; i += WB_ASM_BATCH) {
            constexpr int B = WB_ASM_BATCH;
            double rr[8 * B];
#pragma unroll
            for (int k = 0; k < 8 * B; ++k)
                rr[k] = (i + k / 8 < Q) ? rec[(i + Q * j) * 8 + k] : 0.0;
#pragma unroll
            for (int u = 0; u < B; ++u) {
                if (i + u >= Q) break;
                const double *r = rr + 8 * u;
                const double R0 = r[0] - x0, R1 = r[1] - x1, R2 = r[2] - x2;
                const double r2 = __builtin_fma(R0, R0, __builtin_fma(R1, R1, R2 * R2));
                const double ri = FAST ? rsqrt_f64_newton1(r2) : rsqrt_f64(r2);
                sg = __builtin_fma(ri, r[6], sg);   // sum JxW / r
                sg1 = __builtin_fma(ri, r[7], sg1); // sum xi JxW / r
                const double dn = __builtin_fma(R0, r[3], __builtin_fma(R1, r[4], R2 * r[5]));
                const double thv = (dn * ri) * (ri * ri);
                sh += thv;
                sh1 = __builtin_fma(nodes.phi[i + u], thv, sh1);
            }
        }
        const double sg0 = sg - sg1, sh0 = sh - sh1;  // sum (1 - xi_i) t
        const double eta = nodes.phi[j], ome = 1.0 - eta;
        g0 = __builtin_fma(ome, sg0, g0);
        g1 = __builtin_fma(ome, sg1, g1);
        g2 = __builtin_fma(eta, sg0, g2);
        g3 = __builtin_fma(eta, sg1, g3);
        h0 = __builtin_fma(ome, sh0, h0);
        h1 = __builtin_fma(ome, sh1, h1);
        h2 = __builtin_fma(eta, sh0, h2);
        h3 = __builtin_fma(eta, sh1, h3);
    }
    g[0] = g0;
    g[1] = g1;
    g[2] = g2;
    g[3] = g3;
    h[0] = h0;
    h[1] = h1;
    h[2] = h2;
    h[3] = h3;
}

template <int Q, int TC, int OUT>
__global__ __launch_bounds__(TILE_ROWS) __attribute__((amdgpu_waves_per_eu(WB_ASM_WAVES)))
void k_assemble_tiles(
    const double *__restrict__ xrow, const double *__restrict__ cellq,
    const int *__restrict__ tile_ptr, const int *__restrict__ tile_cell,
    const int *__restrict__ tile_cols, const int *__restrict__ tile_start, int n, int n_tiles,
    size_t ld, PhiTable nodes,
    void *__restrict__ Gv, void *__restrict__ Hv, uint16_t *__restrict__ Glo,
    uint16_t *__restrict__ Hlo, double *__restrict__ hsum_part)
{
    constexpr int NQ = Q * Q;
    static_assert(TC == 16, "tile width");
    typedef double accv __attribute__((ext_vector_type(16)));
    // [column][row] tile in LDS, row stride ACC_LD (68: the remapped row-segment reads of
    // the write-out hit distinct banks); holds the H accumulators, then stages G
    constexpr int ACC_LD = TILE_ROWS + 4;
    __shared__ double lacc[TC * ACC_LD];
    const int lane = threadIdx.x;
    const int rb = blockIdx.x;
    const int tg = blockIdx.y;
    const int row = rb * TILE_ROWS + lane;
    const int rr = row < n ? row : n - 1;
    const double x0 = xrow[3 * rr], x1 = xrow[3 * rr + 1], x2 = xrow[3 * rr + 2];
    double hs_total = 0.0;
    const int t_begin = tg * TILES_PER_WG;
    const int t_end = min(n_tiles, t_begin + TILES_PER_WG);
    for (int tile = t_begin; tile < t_end; ++tile) {
        accv aG = 0.0;
#if WB_ASM_HLDS
#pragma unroll
        for (int c = 0; c < TC; ++c) lacc[c * ACC_LD + lane] = 0.0;
#else
        accv aH = 0.0;
#endif
        // c is wave-uniform: the G vector index becomes an s_set_gpr_idx register move; the
        // H entry is a lane-private LDS word, updated by a no-return LDS add (ds_add_f64)
        auto upd = [&](int c, double g, double h) {
            aG[c & 15] += g;
#if WB_ASM_HLDS
            __hip_atomic_fetch_add(&lacc[(c & 15) * ACC_LD + lane], h, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WAVEFRONT);
#else
            aH[c & 15] += h;
#endif
        };
        double hs = 0.0;
        const int p0 = tile_ptr[tile], p1 = tile_ptr[tile + 1];
        for (int p = p0; p < p1; ++p) {
            const int cell = __builtin_amdgcn_readfirstlane(tile_cell[p]);
            const int cols = __builtin_amdgcn_readfirstlane(tile_cols[p]);
            const double *rec = cellq + (size_t)cell * (NQ * 8);
            double gl[4], hl[4];
            cell_local_values<Q, OUT == 1>(rec, x0, x1, x2, nodes, gl, hl);
            const double g0 = gl[0], g1 = gl[1], g2 = gl[2], g3 = gl[3];
            const double h0 = hl[0], h1 = hl[1], h2 = hl[2], h3 = hl[3];
            const int c0 = (int)(int8_t)(cols & 0xff);
            const int c1 = (int)(int8_t)((cols >> 8) & 0xff);
            const int c2 = (int)(int8_t)((cols >> 16) & 0xff);
            const int c3 = (int)(int8_t)((cols >> 24) & 0xff);
            if (c0 >= 0) {
                upd(c0, g0, h0);
                hs += h0;
            }
            if (c1 >= 0) {
                upd(c1, g1, h1);
                hs += h1;
            }
            if (c2 >= 0) {
                upd(c2, g2, h2);
                hs += h2;
            }
            if (c3 >= 0) {
                upd(c3, g3, h3);
                hs += h3;
            }
        }
        // Write-out through LDS: H (already there), then G parked by each lane; lanes are
        // remapped to (row, column) so every store instruction writes whole row segments.
        const int c_begin = tile_start[tile];
        const int ncols = min(TC, tile_start[tile + 1] - c_begin);
#pragma unroll
        for (int m = 1; m >= 0; --m) {
            const double scale = m == 0 ? INV4PI : -INV4PI;
            __syncthreads();
#if WB_ASM_HLDS
            if (m == 0) {
#pragma unroll
                for (int c = 0; c < TC; ++c) lacc[c * ACC_LD + lane] = aG[c];
            }
#else
#pragma unroll
            for (int c = 0; c < TC; ++c) lacc[c * ACC_LD + lane] = m == 0 ? aG[c] : aH[c];
#endif
            __syncthreads();
            constexpr int RPI = TILE_ROWS / TC;  // rows per store instruction
            const int cl = lane % TC, rsub = lane / TC;
#pragma unroll
            for (int r0 = 0; r0 < TILE_ROWS; r0 += RPI) {
                const int rl = r0 + rsub;
                const int grow = rb * TILE_ROWS + rl;
                if (grow < n && cl < ncols) {
                    const double v = lacc[cl * ACC_LD + rl] * scale;
                    const size_t off = (size_t)grow * ld + c_begin + cl;
                    if (OUT == 0) {
                        ((double *)(m == 0 ? Gv : Hv))[off] = v;
                    } else {
                        float hi;
                        uint16_t lo;
                        split_hilo(v, hi, lo);
                        ((float *)(m == 0 ? Gv : Hv))[off] = hi;
                        (m == 0 ? Glo : Hlo)[off] = lo;
                    }
                }
            }
        }
        hs_total += hs;
    }
    if (row < n) hsum_part[(size_t)tg * n + row] = hs_total * (-INV4PI);
}

template <int Q>
static void launch_tiles_q(int precision, const double *xrow, const double *cellq,
                           const int *tile_ptr, const int *tile_cell, const int *tile_cols,
                           const int *tile_start, int n, int n_tiles, size_t ld,
                           const PhiTable &nodes, void *G, void *H, void *Glo, void *Hlo,
                           double *hsum_part, hipStream_t s)
{
    dim3 grid((n + TILE_ROWS - 1) / TILE_ROWS, (n_tiles + TILES_PER_WG - 1) / TILES_PER_WG);
    if (precision == WB_PREC_FP64)
        hipLaunchKernelGGL((k_assemble_tiles<Q, TILE_COLS, 0>), grid, dim3(TILE_ROWS), 0, s, xrow,
                           cellq, tile_ptr, tile_cell, tile_cols, tile_start, n, n_tiles, ld,
                           nodes, G, H, (uint16_t *)Glo, (uint16_t *)Hlo, hsum_part);
    else
        hipLaunchKernelGGL((k_assemble_tiles<Q, TILE_COLS, 1>), grid, dim3(TILE_ROWS), 0, s, xrow,
                           cellq, tile_ptr, tile_cell, tile_cols, tile_start, n, n_tiles, ld,
                           nodes, G, H, (uint16_t *)Glo, (uint16_t *)Hlo, hsum_part);
}

// nodes.phi[0..q_order-1] = 1-D Gauss nodes on [0,1] of the regular rule.  Tiles of up to
// 16 columns share the 16-wide register tile (tc only shapes the tiles on the host).
void launch_assemble_tiles(int precision, int q_order, int tc, const double *xrow,
                           const double *cellq, const int *tile_ptr, const int *tile_cell,
                           const int *tile_cols, const int *tile_start, int n, int n_tiles,
                           size_t ld, const PhiTable &nodes, void *G, void *H, void *Glo,
                           void *Hlo, double *hsum_part, hipStream_t s)
{
    (void)tc;
#define WB_TILES(QV)                                                                           \
    case QV:                                                                                   \
        launch_tiles_q<QV>(precision, xrow, cellq, tile_ptr, tile_cell, tile_cols, tile_start, \
                           n, n_tiles, ld, nodes, G, H, Glo, Hlo, hsum_part, s);               \
        break;
    switch (q_order) {
        WB_TILES(1)
        WB_TILES(2)
        WB_TILES(3)
        WB_TILES(4)
        WB_TILES(5)
        WB_TILES(6)
    default: break;
    }
#undef WB_TILES
}

// ------------------------------------------------------------------ regular cells, strips
// One wavefront owns 64 collocation rows (lane = row) and one strip task: the lines 0 ..
// L - 1 (grid rows of at most STRIP_W consecutive dofs) of a run of levels of a structured
// patch, swept gap by gap: gap t holds the cells between lines t and t + 1, t = -1 .. L - 1
// (line -1 and line L belong to the neighbouring tasks).  Their records are built from
// cell_int_s, whose vertex order puts the position direction first, so the cell at
// position k of gap t adds its local values 0, 1 to positions k - 1, k of line t ("lo")
// and 2, 3 to positions k - 1, k of line t + 1 ("up").  The position loop is unrolled, so
// every update goes to a fixed register (G) or a fixed LDS offset (H, lane-private words).
// After gap t, line t is complete: it is written once through LDS as whole row segments
// and cleared, and the accumulators of line t + 1 become the "lo" ones of the next gap.
// Every cell is evaluated once per task it touches: once inside a strip, twice on the
// boundary between two strips or two tasks of one strip.
template <int K>
__device__ __forceinline__ void strip_add(double (&lo)[STRIP_W], double (&up)[STRIP_W],
                                          double *hlo, double *hup, const double (&g)[4],
                                          const double (&h)[4], int acc_ld)
{
    if constexpr (K >= 1) {
        lo[K - 1] += g[0];
        up[K - 1] += g[2];
        __hip_atomic_fetch_add(hlo + (K - 1) * acc_ld, h[0], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WAVEFRONT);
        __hip_atomic_fetch_add(hup + (K - 1) * acc_ld, h[2], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
    if constexpr (K < STRIP_W) {
        lo[K] += g[1];
        up[K] += g[3];
        __hip_atomic_fetch_add(hlo + K * acc_ld, h[1], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WAVEFRONT);
        __hip_atomic_fetch_add(hup + K * acc_ld, h[3], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
}

// The cells of one gap, positions K .. STRIP_W (present where cmask has the bit).
template <int Q, int K, bool FAST>
__device__ __forceinline__ void strip_gap_cells(int cmask, const double *grec, double x0,
                                                double x1, double x2, const PhiTable &nodes,
                                                double (&lo)[STRIP_W], double (&up)[STRIP_W],
                                                double *hlo, double *hup, int acc_ld)
{
    if constexpr (K <= STRIP_W) {
        if (cmask & (1 << K)) {
            double cg[4], ch[4];
            cell_local_values<Q, FAST>(grec + K * (Q * Q * 8), x0, x1, x2, nodes, cg, ch);
            strip_add<K>(lo, up, hlo, hup, cg, ch, acc_ld);
        }
        strip_gap_cells<Q, K + 1, FAST>(cmask, grec, x0, x1, x2, nodes, lo, up, hlo, hup, acc_ld);
    }
}

template <int Q, int OUT>
__global__ __launch_bounds__(TILE_ROWS) __attribute__((amdgpu_waves_per_eu(WB_ASM_WAVES)))
void k_assemble_strips(const double *__restrict__ xrow, const double *__restrict__ recs,
                       const int2 *__restrict__ tasks, const int4 *__restrict__ gaps, int n,
                       size_t ld, PhiTable nodes, void *__restrict__ Gv, void *__restrict__ Hv,
                       uint16_t *__restrict__ Glo, uint16_t *__restrict__ Hlo,
                       double *__restrict__ hsum_part)
{
    constexpr int NQ = Q * Q;
    constexpr int KS = STRIP_W + 1;  // cell positions per gap
    // H accumulators: two halves of STRIP_W lane-private columns [slot][lane], row stride
    // ACC_LD; the half of line t + 1 follows the half of line t in turn
    constexpr int ACC_LD = TILE_ROWS + 4;
    __shared__ double lacc[2 * STRIP_W * ACC_LD];
    const int lane = threadIdx.x;
    const int rb = blockIdx.x;
    const int task = blockIdx.y;
    const int row = rb * TILE_ROWS + lane;
    const int rr = row < n ? row : n - 1;
    const double x0 = xrow[3 * rr], x1 = xrow[3 * rr + 1], x2 = xrow[3 * rr + 2];
    const int2 T = tasks[task];
    const int g0 = __builtin_amdgcn_readfirstlane(T.x);
    const int g1 = g0 + __builtin_amdgcn_readfirstlane(T.y);
    double lo[STRIP_W], up[STRIP_W];
#pragma unroll
    for (int q = 0; q < STRIP_W; ++q) {
        lo[q] = up[q] = 0.0;
        lacc[q * ACC_LD + lane] = 0.0;
        lacc[(STRIP_W + q) * ACC_LD + lane] = 0.0;
    }
    int hl = 0;  // LDS half of the lo line
    double hs = 0.0;
    // lane -> (position p, row offset rsub) in the write-out: each store instruction writes
    // STRIP_W-column row segments of 64 / STRIP_W rows
    constexpr int RPI = TILE_ROWS / STRIP_W;
    const int wp = lane % STRIP_W, wsub = lane / STRIP_W;
    for (int g = g0; g < g1; ++g) {
        const int4 gi = gaps[g];
        const int cmask = __builtin_amdgcn_readfirstlane(gi.x);
        const int col0 = __builtin_amdgcn_readfirstlane(gi.y);
        const int lmask = __builtin_amdgcn_readfirstlane(gi.z);
        double *hlo = lacc + hl * (STRIP_W * ACC_LD) + lane;
        double *hup = lacc + (1 - hl) * (STRIP_W * ACC_LD) + lane;
        const double *grec = recs + (size_t)g * KS * (NQ * 8);
        strip_gap_cells<Q, 0, OUT == 1>(cmask, grec, x0, x1, x2, nodes, lo, up, hlo, hup, ACC_LD);
        // line t is complete: row sums, then H and G out through LDS, then clear
        double *hbase = lacc + hl * (STRIP_W * ACC_LD);
#pragma unroll
        for (int p = 0; p < STRIP_W; ++p)
            if (lmask & (1 << p)) hs += hbase[p * ACC_LD + lane];
        if (lmask) {
            const bool present = (lmask >> wp) & 1;
            const int col = col0 + __builtin_popcount(lmask & ((1 << wp) - 1));
#pragma unroll
            for (int m = 1; m >= 0; --m) {  // H, then G
                const double scale = m == 0 ? INV4PI : -INV4PI;
                __syncthreads();
                if (m == 0) {
#pragma unroll
                    for (int q = 0; q < STRIP_W; ++q) hbase[q * ACC_LD + lane] = lo[q];
                    __syncthreads();
                }
#pragma unroll
                for (int r0 = 0; r0 < TILE_ROWS; r0 += RPI) {
                    const int rl = r0 + wsub;
                    const int grow = rb * TILE_ROWS + rl;
                    if (grow < n && present) {
                        const double v = hbase[wp * ACC_LD + rl] * scale;
                        const size_t off = (size_t)grow * ld + col;
                        if (OUT == 0) {
                            ((double *)(m == 0 ? Gv : Hv))[off] = v;
                        } else {
                            float hi;
                            uint16_t lo16;
                            split_hilo(v, hi, lo16);
                            ((float *)(m == 0 ? Gv : Hv))[off] = hi;
                            (m == 0 ? Glo : Hlo)[off] = lo16;
                        }
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < STRIP_W; ++q) {
            hbase[q * ACC_LD + lane] = 0.0;
            lo[q] = up[q];
            up[q] = 0.0;
        }
        hl ^= 1;
    }
    if (row < n) hsum_part[(size_t)task * n + row] = hs * (-INV4PI);
}

void launch_assemble_strips(int precision, int q_order, const double *xrow, const double *recs,
                            const int *tasks, const int *gaps, int n, int n_tasks, size_t ld,
                            const PhiTable &nodes, void *G, void *H, void *Glo, void *Hlo,
                            double *hsum_part, hipStream_t s)
{
    if (n_tasks <= 0) return;
    dim3 grid((n + TILE_ROWS - 1) / TILE_ROWS, n_tasks);
#define WB_STRIPS(QV)                                                                          \
    case QV:                                                                                   \
        if (precision == WB_PREC_FP64)                                                         \
            hipLaunchKernelGGL((k_assemble_strips<QV, 0>), grid, dim3(TILE_ROWS), 0, s, xrow,  \
                               recs, (const int2 *)tasks, (const int4 *)gaps, n, ld, nodes, G, \
                               H, (uint16_t *)Glo, (uint16_t *)Hlo, hsum_part);                \
        else                                                                                   \
            hipLaunchKernelGGL((k_assemble_strips<QV, 1>), grid, dim3(TILE_ROWS), 0, s, xrow,  \
                               recs, (const int2 *)tasks, (const int4 *)gaps, n, ld, nodes, G, \
                               H, (uint16_t *)Glo, (uint16_t *)Hlo, hsum_part);                \
        break;
    switch (q_order) {
        WB_STRIPS(1)
        WB_STRIPS(2)
        WB_STRIPS(3)
        WB_STRIPS(4)
        WB_STRIPS(5)
        WB_STRIPS(6)
    default: break;
    }
#undef WB_STRIPS
}

// ------------------------------------------------------------------ singular cells
// One thread per row: for each cell touching x_i (first coincident vertex j*), evaluate
// the QGaussOneOverR rule of vertex j* with full FEValues geometry at every point
// (source/bem_problem.cc:475-525) and add the 4 local values into row i, less the
// regular-rule values the regular kernels added for the same pair (they evaluate every
// cell with QGauss, which keeps the singular test out of their inner loop).
template <int OUT>
__global__ __launch_bounds__(64) void k_assemble_singular(
    const double *__restrict__ X, const int *__restrict__ cell_int,
    const int8_t *__restrict__ orient, const int *__restrict__ sing_ptr,
    const int *__restrict__ sing_cell, const int *__restrict__ sing_vtx, int n, int ns,
    const double *__restrict__ rule, const double *__restrict__ cellq, int nq,
    const double *__restrict__ qpts, size_t ld, void *__restrict__ Gv, void *__restrict__ Hv,
    uint16_t *__restrict__ Glo, uint16_t *__restrict__ Hlo, double *__restrict__ hsum_sing)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x0 = X[3 * i], x1 = X[3 * i + 1], x2 = X[3 * i + 2];
    double hs = 0.0;
    for (int p = sing_ptr[i]; p < sing_ptr[i + 1]; ++p) {
        const int c = sing_cell[p], vtx = sing_vtx[p];
        int vid[4];
        double v[4][3];
        for (int k = 0; k < 4; ++k) {
            vid[k] = cell_int[4 * c + k];
            for (int d = 0; d < 3; ++d) v[k][d] = X[3 * vid[k] + d];
        }
        const double sgn = orient[c] < 0 ? -1.0 : 1.0;
        const double *ru = rule + (size_t)vtx * ns * 3;
        double lg[4] = {0, 0, 0, 0}, lh[4] = {0, 0, 0, 0};
        for (int q = 0; q < ns; ++q) {
            const double xi = ru[3 * q], eta = ru[3 * q + 1], w = ru[3 * q + 2];
            const double phi[4] = {(1 - xi) * (1 - eta), xi * (1 - eta), (1 - xi) * eta, xi * eta};
            const double dxi[4] = {-(1 - eta), (1 - eta), -eta, eta};
            const double deta[4] = {-(1 - xi), -xi, (1 - xi), xi};
            double y[3] = {0, 0, 0}, a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
            for (int k = 0; k < 4; ++k)
                for (int d = 0; d < 3; ++d) {
                    y[d] += phi[k] * v[k][d];
                    a[d] += dxi[k] * v[k][d];
                    b[d] += deta[k] * v[k][d];
                }
            const double g00 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
            const double g01 = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
            const double g11 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
            const double jxw = sqrt(g00 * g11 - g01 * g01) * w;
            double nn[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                            a[0] * b[1] - a[1] * b[0]};
            const double nl = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
            const double R0 = y[0] - x0, R1 = y[1] - x1, R2 = y[2] - x2;
            const double r2 = R0 * R0 + R1 * R1 + R2 * R2;
            const double ri = rsqrt_f64(r2);
            const double dn = (R0 * nn[0] + R1 * nn[1] + R2 * nn[2]) * (sgn / nl);
            const double tg = ri * jxw;
            const double th = dn * ri * ri * ri * jxw;
            for (int k = 0; k < 4; ++k) {
                lg[k] = __builtin_fma(tg, phi[k], lg[k]);
                lh[k] = __builtin_fma(th, phi[k], lh[k]);
            }
        }
        // the regular-rule part the regular kernels added for this pair (same records,
        // same kernel; it cancels up to rounding)
        const double *rec = cellq + (size_t)c * nq * 8;
        for (int q = 0; q < nq; ++q) {
            const double *r = rec + 8 * q;
            const double xi = qpts[2 * q], eta = qpts[2 * q + 1];
            const double phi[4] = {(1 - xi) * (1 - eta), xi * (1 - eta), (1 - xi) * eta, xi * eta};
            const double R0 = r[0] - x0, R1 = r[1] - x1, R2 = r[2] - x2;
            const double ri = rsqrt_f64(__builtin_fma(R0, R0, __builtin_fma(R1, R1, R2 * R2)));
            const double dn = __builtin_fma(R0, r[3], __builtin_fma(R1, r[4], R2 * r[5]));
            const double tg = ri * r[6], th = (dn * ri) * (ri * ri);
            for (int k = 0; k < 4; ++k) {
                lg[k] = __builtin_fma(-tg, phi[k], lg[k]);
                lh[k] = __builtin_fma(-th, phi[k], lh[k]);
            }
        }
        for (int k = 0; k < 4; ++k) {
            const double gadd = lg[k] * INV4PI, hadd = lh[k] * (-INV4PI);
            const size_t off = (size_t)i * ld + vid[k];
            if (OUT == 0) {
                ((double *)Gv)[off] += gadd;
                ((double *)Hv)[off] += hadd;
            } else {
                float gh, hh;
                uint16_t gl, hl;
                const double gv = join_hilo(((float *)Gv)[off], Glo[off]) + gadd;
                const double hv = join_hilo(((float *)Hv)[off], Hlo[off]) + hadd;
                split_hilo(gv, gh, gl);
                split_hilo(hv, hh, hl);
                ((float *)Gv)[off] = gh;
                ((float *)Hv)[off] = hh;
                Glo[off] = gl;
                Hlo[off] = hl;
            }
            hs += hadd;
        }
    }
    hsum_sing[i] = hs;
}

void launch_assemble_singular(int precision, const double *X, const int *cell_int,
                              const int8_t *orient, const int *sing_ptr, const int *sing_cell,
                              const int *sing_vtx, int n, int ns, const double *rule,
                              const double *cellq, int nq, const double *qpts, size_t ld,
                              void *G, void *H, void *Glo, void *Hlo, double *hsum_sing,
                              hipStream_t s)
{
    dim3 grid((n + 63) / 64);
    if (precision == WB_PREC_FP64)
        hipLaunchKernelGGL(k_assemble_singular<0>, grid, dim3(64), 0, s, X, cell_int, orient,
                           sing_ptr, sing_cell, sing_vtx, n, ns, rule, cellq, nq, qpts, ld, G, H,
                           (uint16_t *)Glo, (uint16_t *)Hlo, hsum_sing);
    else
        hipLaunchKernelGGL(k_assemble_singular<1>, grid, dim3(64), 0, s, X, cell_int, orient,
                           sing_ptr, sing_cell, sing_vtx, n, ns, rule, cellq, nq, qpts, ld, G, H,
                           (uint16_t *)Glo, (uint16_t *)Hlo, hsum_sing);
}

// alpha_i = -(sum over tile groups of the regular H row sums + singular H row sum)
// alpha_i = -(sum over tile groups of the partial row sums + singular part).  A block
// owns 64 rows; its 4 wavefronts sum interleaved quarters of the partials and the quarters
// are added in a fixed order, so the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_alpha_reduce(const double *__restrict__ part,
                                                      int n_parts,
                                                      const double *__restrict__ sing, int n,
                                                      double *__restrict__ alpha)
{
    __shared__ double q[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    double s = 0.0;
    if (i < n)
        for (int p = wave; p < n_parts; p += 4) s += part[(size_t)p * n + i];
    q[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && i < n) alpha[i] = -(((q[0][lane] + q[1][lane]) + (q[2][lane] + q[3][lane])) + sing[i]);
}

void launch_alpha_reduce(const double *hsum_part, int n_parts, const double *hsum_sing, int n,
                         double *alpha, hipStream_t s)
{
    hipLaunchKernelGGL(k_alpha_reduce, dim3((n + 63) / 64), dim3(256), 0, s, hsum_part, n_parts,
                       hsum_sing, n, alpha);
}

}  // namespace wb
