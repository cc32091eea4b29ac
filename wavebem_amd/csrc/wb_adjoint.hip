// wb_adjoint.hip -- transposed operator and the backward of the assembly on gfx950.
//
// Adjoint of the discrete BEM solve (no counterpart in the reference, whose Jacobians
// come from Sacado over the same residual, source/free_surface.cc:4918-5053 and
// :6099-6109).  For a functional J(u) of the solution of A_c u = b (the constrained
// system of source/bem_problem.cc:821-895), dJ/dtheta = dJ/dtheta|_u - lambda^T dR/dtheta
// with A_c^T lambda = dJ/du.  This file holds:
//   * the transposed column-selected product A_sel^T x (kernel k_gemvT, FP64 entries, the
//     FP32 hi part alone, or hi + bf16 lo accumulated in FP64): every
//     workgroup owns a strip of 128 columns and a chunk of 512 rows; lanes read whole row
//     segments (coalesced, like the forward product), four rows per step, and keep their
//     two columns' sums in registers; the chunk partials are added in a fixed order;
//   * the transposed constraint lines (k_constraint_T, a column-wise gather built on the
//     host, so there are no atomics);
//   * the backward of assemble_system with respect to the support points: for row i with
//     weight w_i, R_i = sum_j H_ij (phi_j - phi_i) - sum_j G_ij psi_j, and each
//     (row, cell, point) term s = w_i/(4 pi) [ -(Phi_q - phi_i) (R.N) / r^3 - Psi_q J / r ]
//     (R = y_q - x_i, N = n_q JxW_q, J = JxW_q, Phi/Psi the Q1 interpolants) has
//       ds/dR = w/(4 pi) [ t1 N + t2 R ],  t1 = -(Phi - phi_i)/r^3,
//               t2 = 3 (Phi - phi_i) (R.N)/r^5 + Psi J / r^3,
//       ds/dN = w/(4 pi) t1 R,  ds/dJ = -w/(4 pi) Psi / r,
//     with d/dx_i = -d/dR and d/dy_q = +d/dR.  Two passes share the arithmetic:
//     k_vjp_rows (lane = row, sums over points: the x_i part) and k_vjp_points
//     (lane = cell point, sums over rows: the y, N, J parts), then k_vjp_map chains
//     y, N = o w (a x b), J = w |a x b| to the four vertices.  Singular (row, cell) pairs
//     use the QGaussOneOverR points in k_vjp_singular, one thread per row.
#include "wb_device.h"
#include "wb_internal.h"

// Wave-uniform records (point records of a cell, row records) are read by scalar loads,
// which return out of order: each use waits for every outstanding scalar load.  Records
// are therefore fetched in batches, one wait per batch.
#ifndef WB_ADJ_BATCH
#define WB_ADJ_BATCH 4  // points (rows) per batch in k_vjp_rows and k_vjp_points
#endif
#ifndef WB_JVP_BATCH
#define WB_JVP_BATCH 2  // points per batch in k_jvp_rows (18 doubles per point)
#endif

namespace wb {

namespace {

typedef double dvec2 __attribute__((ext_vector_type(2)));

constexpr int GT_WAVES = 4;
constexpr int GT_COLS = 128;  // columns per workgroup: 64 lanes x 2

// The column pair (col, col + 1) of row offset off, as FP64: MODE 0 FP64 entries, MODE 1
// the FP32 hi part, MODE 2 hi + bf16 lo.  M0/M1 (and L0/L1) are the pair's matrices.
template <int MODE>
__device__ __forceinline__ void load_pair(const void *M0, const void *M1, const uint16_t *L0,
                                          const uint16_t *L1, size_t off, bool same, double &v0,
                                          double &v1)
{
    if (MODE == 0) {
        if (same) {
            const dvec2 v = __builtin_nontemporal_load((const dvec2 *)((const double *)M0 + off));
            v0 = v.x;
            v1 = v.y;
        } else {
            v0 = ((const double *)M0)[off];
            v1 = ((const double *)M1)[off + 1];
        }
    } else {
        float h0, h1;
        if (same) {
            const float2 h = *(const float2 *)((const float *)M0 + off);
            h0 = h.x;
            h1 = h.y;
        } else {
            h0 = ((const float *)M0)[off];
            h1 = ((const float *)M1)[off + 1];
        }
        v0 = h0;
        v1 = h1;
        if (MODE == 2) {
            uint16_t l0, l1;
            if (same) {
                const ushort2 l = *(const ushort2 *)(L0 + off);
                l0 = l.x;
                l1 = l.y;
            } else {
                l0 = L0[off];
                l1 = L1[off + 1];
            }
            v0 += (double)__uint_as_float((uint32_t)l0 << 16);
            v1 += (double)__uint_as_float((uint32_t)l1 << 16);
        }
    }
}

// part[chunk][col] = sum over the chunk's unconstrained rows i of
// sel(col) * A_sel(col)(i, col) * x[i], accumulated in FP64 for every storage MODE (x of
// type XT: double, or float for the FP32 inner solves).  Rows are taken in order
// (constrained rows read with weight 0), four per wavefront step, so the row addresses are
// affine and the loads of a step are independent.  Columns come in pairs, so a pair at an
// odd-aligned column of FP32 or bf16 storage still loads 8 and 4 bytes.
template <int MODE, typename XT>
__global__ __launch_bounds__(64 * GT_WAVES) void k_gemvT(
    const void *__restrict__ G, const void *__restrict__ H, const uint16_t *__restrict__ Glo,
    const uint16_t *__restrict__ Hlo, size_t ld, int n, const uint8_t *__restrict__ gmask,
    int flip, double s_sel, double s_uns, const XT *__restrict__ x,
    const uint8_t *__restrict__ constrained, int rows_per_chunk, double *__restrict__ part)
{
    __shared__ double red[GT_WAVES][64][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = blockIdx.x * GT_COLS + 2 * lane;  // this lane's column pair
    const bool live = col < n;
    unsigned b0 = 0, b1 = 0;
    if (live) {
        const unsigned gm = (unsigned)gmask[col >> 2] ^ (flip ? 0xfu : 0u);
        b0 = (gm >> (col & 3)) & 1u;
        b1 = (gm >> ((col + 1) & 3)) & 1u;
    }
    const void *M0 = b0 ? H : G, *M1 = b1 ? H : G;
    const uint16_t *L0 = b0 ? Hlo : Glo, *L1 = b1 ? Hlo : Glo;
    const bool same = b0 == b1;
    double a0 = 0.0, a1 = 0.0;
    const int i0 = blockIdx.y * rows_per_chunk;
    const int i1 = min(n, i0 + rows_per_chunk);
    constexpr int U = 4;  // rows per wavefront step
    auto xw = [&](int i) { return constrained[i] ? 0.0 : (double)x[i]; };
    if (live) {
        const int nfull = (i1 - i0) / U;  // whole groups of U rows; group g to wave g % 4
        for (int g = wave; g < nfull; g += GT_WAVES) {
            const int i = i0 + g * U;
            double xi[U], v0[U], v1[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xi[u] = xw(i + u);
#pragma unroll
            for (int u = 0; u < U; ++u)
                load_pair<MODE>(M0, M1, L0, L1, (size_t)(i + u) * ld + col, same, v0[u], v1[u]);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                a0 += v0[u] * xi[u];
                a1 += v1[u] * xi[u];
            }
        }
        if (wave == 0)  // the last (i1 - i0) % U rows
            for (int i = i0 + nfull * U; i < i1; ++i) {
                const double xv = xw(i);
                double v0, v1;
                load_pair<MODE>(M0, M1, L0, L1, (size_t)i * ld + col, false, v0, v1);
                a0 += v0 * xv;
                a1 += v1 * xv;
            }
    }
    red[wave][lane][0] = a0 * (b0 ? s_sel : s_uns);
    red[wave][lane][1] = a1 * (b1 ? s_sel : s_uns);
    __syncthreads();
    if (wave == 0 && live) {
        double t0 = red[0][lane][0], t1 = red[0][lane][1];
#pragma unroll
        for (int w = 1; w < GT_WAVES; ++w) {
            t0 += red[w][lane][0];
            t1 += red[w][lane][1];
        }
        double *pp = part + (size_t)blockIdx.y * ld + col;
        pp[0] = t0;
        if (col + 1 < n) pp[1] = t1;
    }
}

// y[j] = sum_chunks part[c][j] + (row j unconstrained ? diag[j] x[j] : 0)
template <typename XT>
__global__ void k_gemvT_reduce(const double *__restrict__ part, int n_chunks, size_t ld, int n,
                               const double *__restrict__ diag,
                               const uint8_t *__restrict__ constrained,
                               const XT *__restrict__ x, XT *__restrict__ y)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double s = 0.0;
    for (int c = 0; c < n_chunks; ++c) s += part[(size_t)c * ld + j];
    if (diag && !(constrained && constrained[j])) s += diag[j] * (double)x[j];
    y[j] = (XT)s;
}

// y[tgt[t]] += sum_p coef[p] x[src[p]]  (targets distinct)
template <typename T>
__global__ void k_constraint_T(int n_targets, const int *__restrict__ tgt,
                               const int *__restrict__ ptr, const int *__restrict__ src,
                               const double *__restrict__ coef, const T *__restrict__ x,
                               T *__restrict__ y)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_targets) return;
    double s = y[tgt[t]];
    for (int p = ptr[t]; p < ptr[t + 1]; ++p) s += coef[p] * (double)x[src[p]];
    y[tgt[t]] = (T)s;
}

// Phi and Psi (Q1 interpolants of phi and psi) at every regular point of every cell
__global__ void k_cell_state(int nc, int nq, const int *__restrict__ cell_int,
                             const double *__restrict__ qpts, const double *__restrict__ phi,
                             const double *__restrict__ psi, double *__restrict__ cstate)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nc * nq) return;
    const int c = gid / nq, q = gid % nq;
    double sh[4], dx[4], de[4];
    q1_shape(qpts[2 * q], qpts[2 * q + 1], sh, dx, de);
    double P = 0.0, S = 0.0;
    for (int k = 0; k < 4; ++k) {
        const int v = cell_int[4 * c + k];
        P += sh[k] * phi[v];
        S += sh[k] * psi[v];
    }
    cstate[2 * (size_t)gid] = P;
    cstate[2 * (size_t)gid + 1] = S;
}

// Pass 1: part[chunk][row][3] = -w_i/(4 pi) sum over the chunk's regular (row, cell) pairs
// of ds/dR (the x_i part of the backward).  Lane = row; cell data are wave-uniform.
template <int Q>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_vjp_rows(
    const double *__restrict__ xrow, const int *__restrict__ row_cls,
    const double *__restrict__ wrow, const double *__restrict__ phi,
    const double *__restrict__ cellq, const double *__restrict__ cstate,
    const int4 *__restrict__ cell_cls, int nc, int cells_per_chunk, int n,
    double *__restrict__ part)
{
    constexpr int NQ = Q * Q;
    const int lane = threadIdx.x;
    const int row = blockIdx.x * 64 + lane;
    const int rr = row < n ? row : n - 1;
    const double x0 = xrow[3 * rr], x1 = xrow[3 * rr + 1], x2 = xrow[3 * rr + 2];
    const int rc = row_cls[rr];
    const double wi = row < n ? wrow[rr] : 0.0, phi_i = phi[rr];
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    const int c0 = blockIdx.y * cells_per_chunk;
    const int c1 = min(nc, c0 + cells_per_chunk);
    if (__ballot(wi != 0.0) != 0) {
        for (int c = c0; c < c1; ++c) {
            const int4 cc = cell_cls[c];
            const double *rec = cellq + (size_t)c * (NQ * 8);
            const double *st = cstate + (size_t)c * (NQ * 2);
            double g0 = 0.0, g1 = 0.0, g2 = 0.0;
            constexpr int B = WB_ADJ_BATCH;
#pragma unroll
            for (int q0 = 0; q0 < NQ; q0 += B) {
                double rb[8 * B], sb[2 * B];
#pragma unroll
                for (int k = 0; k < 8 * B; ++k) rb[k] = (q0 + k / 8 < NQ) ? rec[8 * q0 + k] : 0.0;
#pragma unroll
                for (int k = 0; k < 2 * B; ++k) sb[k] = (q0 + k / 2 < NQ) ? st[2 * q0 + k] : 0.0;
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    if (q0 + u >= NQ) break;
                    const double *r = rb + 8 * u;
                    const double R0 = r[0] - x0, R1 = r[1] - x1, R2 = r[2] - x2;
                    const double r2 = __builtin_fma(R0, R0, __builtin_fma(R1, R1, R2 * R2));
                    const double ri = rsqrt_f64(r2);
                    const double ri2 = ri * ri, ri3 = ri2 * ri, ri5 = ri3 * ri2;
                    const double RN = __builtin_fma(R0, r[3], __builtin_fma(R1, r[4], R2 * r[5]));
                    const double D = sb[2 * u] - phi_i;
                    const double t1 = -D * ri3;
                    const double t2 = __builtin_fma(3.0 * D * RN, ri5, sb[2 * u + 1] * r[6] * ri3);
                    g0 = __builtin_fma(t1, r[3], __builtin_fma(t2, R0, g0));
                    g1 = __builtin_fma(t1, r[4], __builtin_fma(t2, R1, g1));
                    g2 = __builtin_fma(t1, r[5], __builtin_fma(t2, R2, g2));
                }
            }
            const bool sing = (rc == cc.x) | (rc == cc.y) | (rc == cc.z) | (rc == cc.w);
            if (!sing) {
                acc0 += g0;
                acc1 += g1;
                acc2 += g2;
            }
        }
    }
    if (row < n) {
        double *o = part + ((size_t)blockIdx.y * n + row) * 3;
        const double s = -wi * INV4PI;
        o[0] = s * acc0;
        o[1] = s * acc1;
        o[2] = s * acc2;
    }
}

// Pass 2: part[chunk][cell][q][7] = sum over the chunk's rows i (regular pairs only) of
// w_i/(4 pi) (ds/dR (3), ds/dN (3), ds/dJ).  Lane = (cell, point); row data are
// wave-uniform, packed as rowpack[i] = {x, y, z, w_i, phi_i, class}.
template <int Q>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_vjp_points(
    const double *__restrict__ rowpack, int n, int rows_per_chunk,
    const double *__restrict__ cellq, const double *__restrict__ cstate,
    const int4 *__restrict__ cell_cls, int nc, double *__restrict__ part)
{
    constexpr int NQ = Q * Q;
    constexpr int CPW = 64 / NQ;  // cells per wavefront
    const int lane = threadIdx.x;
    const int c = blockIdx.x * CPW + lane / NQ, q = lane % NQ;
    const bool live = lane < CPW * NQ && c < nc;
    const int cs = live ? c : 0;
    const double *r = cellq + ((size_t)cs * NQ + q) * 8;
    const double y0 = r[0], y1 = r[1], y2 = r[2], N0 = r[3], N1 = r[4], N2 = r[5], J = r[6];
    const double P = cstate[2 * ((size_t)cs * NQ + q)], S = cstate[2 * ((size_t)cs * NQ + q) + 1];
    const int4 cc = cell_cls[cs];
    const double SJ = S * J;
    double ay0 = 0, ay1 = 0, ay2 = 0, aN0 = 0, aN1 = 0, aN2 = 0, aJ = 0;
    const int i0 = blockIdx.y * rows_per_chunk;
    const int i1 = min(n, i0 + rows_per_chunk);
    constexpr int B = WB_ADJ_BATCH;
    for (int ib = i0; ib < i1; ib += B) {
        double pb[8 * B];
#pragma unroll
        for (int k = 0; k < 8 * B; ++k)
            pb[k] = (ib + k / 8 < i1) ? rowpack[8 * (size_t)ib + k] : 0.0;
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const double *rp = pb + 8 * u;
            const double wi = rp[3];  // 0 for rows past the chunk (padding above)
            if (wi == 0.0) continue;
            const int ci = (int)rp[5];
            const double R0 = y0 - rp[0], R1 = y1 - rp[1], R2 = y2 - rp[2];
            const double r2 = __builtin_fma(R0, R0, __builtin_fma(R1, R1, R2 * R2));
            const double ri = rsqrt_f64(r2);
            const double ri2 = ri * ri, ri3 = ri2 * ri, ri5 = ri3 * ri2;
            const double RN = __builtin_fma(R0, N0, __builtin_fma(R1, N1, R2 * N2));
            const double D = P - rp[4];
            const double t1 = -D * ri3;
            const double t2 = __builtin_fma(3.0 * D * RN, ri5, SJ * ri3);
            const bool sing = (ci == cc.x) | (ci == cc.y) | (ci == cc.z) | (ci == cc.w);
            const double ws = sing ? 0.0 : wi;
            const double u1 = ws * t1, u2 = ws * t2;
            ay0 = __builtin_fma(u1, N0, __builtin_fma(u2, R0, ay0));
            ay1 = __builtin_fma(u1, N1, __builtin_fma(u2, R1, ay1));
            ay2 = __builtin_fma(u1, N2, __builtin_fma(u2, R2, ay2));
            aN0 = __builtin_fma(u1, R0, aN0);
            aN1 = __builtin_fma(u1, R1, aN1);
            aN2 = __builtin_fma(u1, R2, aN2);
            aJ = __builtin_fma(-ws * S, ri, aJ);
        }
    }
    if (live) {
        double *o = part + (((size_t)blockIdx.y * nc + c) * NQ + q) * 7;
        o[0] = ay0 * INV4PI;
        o[1] = ay1 * INV4PI;
        o[2] = ay2 * INV4PI;
        o[3] = aN0 * INV4PI;
        o[4] = aN1 * INV4PI;
        o[5] = aN2 * INV4PI;
        o[6] = aJ * INV4PI;
    }
}

// Adjoints of (y, N, J) at a point of the Q1 map with vertices v, reference point
// (xi, eta), weight w and orientation sign so, added into vb[4][3]:
// y = sum sh_k v_k, a = sum dxi_k v_k, b = sum deta_k v_k, N = so w (a x b), J = w |a x b|.
__device__ __forceinline__ void chain_point(const double (&v)[4][3], double xi, double eta,
                                            double w, double so, const double *yb,
                                            const double *Nb, double Jb, double (&vb)[4][3])
{
    double sh[4], dx[4], de[4];
    q1_shape(xi, eta, sh, dx, de);
    double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) {
            a[d] += dx[k] * v[k][d];
            b[d] += de[k] * v[k][d];
        }
    const double cx[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                          a[0] * b[1] - a[1] * b[0]};
    const double il = 1.0 / sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
    const double m[3] = {cx[0] * il, cx[1] * il, cx[2] * il};
    const double sN = so * w, sJ = w * Jb;
    const double ga[3] = {sN * (b[1] * Nb[2] - b[2] * Nb[1]) + sJ * (b[1] * m[2] - b[2] * m[1]),
                          sN * (b[2] * Nb[0] - b[0] * Nb[2]) + sJ * (b[2] * m[0] - b[0] * m[2]),
                          sN * (b[0] * Nb[1] - b[1] * Nb[0]) + sJ * (b[0] * m[1] - b[1] * m[0])};
    const double gb[3] = {sN * (Nb[1] * a[2] - Nb[2] * a[1]) + sJ * (m[1] * a[2] - m[2] * a[1]),
                          sN * (Nb[2] * a[0] - Nb[0] * a[2]) + sJ * (m[2] * a[0] - m[0] * a[2]),
                          sN * (Nb[0] * a[1] - Nb[1] * a[0]) + sJ * (m[0] * a[1] - m[1] * a[0])};
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d)
            vb[k][d] += sh[k] * yb[d] + dx[k] * ga[d] + de[k] * gb[d];
}

// vbar[c][k][3]: the regular-point adjoints of cell c (pass-2 chunks added in order)
// chained to its four vertices.  One thread per cell.
__global__ void k_vjp_map(int nc, int nq, int n_chunks, const double *__restrict__ part,
                          const double *__restrict__ X, const int *__restrict__ cell_int,
                          const int8_t *__restrict__ orient, const double *__restrict__ qpts,
                          const double *__restrict__ qw, double *__restrict__ vbar)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double v[4][3];
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) v[k][d] = X[3 * cell_int[4 * c + k] + d];
    const double so = orient[c] < 0 ? -1.0 : 1.0;
    double vb[4][3] = {};
    for (int q = 0; q < nq; ++q) {
        double g[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int ch = 0; ch < n_chunks; ++ch) {
            const double *p = part + (((size_t)ch * nc + c) * nq + q) * 7;
            for (int e = 0; e < 7; ++e) g[e] += p[e];
        }
        chain_point(v, qpts[2 * q], qpts[2 * q + 1], qw[q], so, g, g + 3, g[6], vb);
    }
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) vbar[(size_t)c * 12 + 3 * k + d] = vb[k][d];
}

// Singular pairs: one thread per row i; xs[i][3] gets the x_i part, spart[p][4][3] the
// vertex part of singular pair p (CSR position).
__global__ __launch_bounds__(64) void k_vjp_singular(
    const double *__restrict__ X, const int *__restrict__ cell_int,
    const int8_t *__restrict__ orient, const int *__restrict__ sing_ptr,
    const int *__restrict__ sing_cell, const int *__restrict__ sing_vtx, int n, int ns,
    const double *__restrict__ rule, const double *__restrict__ wrow,
    const double *__restrict__ phi, const double *__restrict__ psi, double *__restrict__ xs,
    double *__restrict__ spart)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x0 = X[3 * i], x1 = X[3 * i + 1], x2 = X[3 * i + 2];
    const double wi = wrow[i] * INV4PI, phi_i = phi[i];
    double gx[3] = {0, 0, 0};
    for (int p = sing_ptr[i]; p < sing_ptr[i + 1]; ++p) {
        const int c = sing_cell[p], vtx = sing_vtx[p];
        double v[4][3], pv[4], sv[4];
        for (int k = 0; k < 4; ++k) {
            const int id = cell_int[4 * c + k];
            for (int d = 0; d < 3; ++d) v[k][d] = X[3 * id + d];
            pv[k] = phi[id];
            sv[k] = psi[id];
        }
        const double so = orient[c] < 0 ? -1.0 : 1.0;
        const double *ru = rule + (size_t)vtx * ns * 3;
        double vb[4][3] = {};
        if (wi != 0.0) {
            for (int q = 0; q < ns; ++q) {
                const double xi = ru[3 * q], eta = ru[3 * q + 1], w = ru[3 * q + 2];
                double sh[4], dx[4], de[4];
                q1_shape(xi, eta, sh, dx, de);
                double y[3] = {0, 0, 0}, a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
                double P = 0.0, S = 0.0;
                for (int k = 0; k < 4; ++k) {
                    for (int d = 0; d < 3; ++d) {
                        y[d] += sh[k] * v[k][d];
                        a[d] += dx[k] * v[k][d];
                        b[d] += de[k] * v[k][d];
                    }
                    P += sh[k] * pv[k];
                    S += sh[k] * sv[k];
                }
                const double cx[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                                      a[0] * b[1] - a[1] * b[0]};
                const double cl = sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
                const double N[3] = {so * w * cx[0], so * w * cx[1], so * w * cx[2]};
                const double J = w * cl;
                const double R[3] = {y[0] - x0, y[1] - x1, y[2] - x2};
                const double ri = rsqrt_f64(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
                const double ri3 = ri * ri * ri, ri5 = ri3 * ri * ri;
                const double RN = R[0] * N[0] + R[1] * N[1] + R[2] * N[2];
                const double D = P - phi_i;
                const double t1 = -D * ri3, t2 = 3.0 * D * RN * ri5 + S * J * ri3;
                double yb[3], Nb[3];
                for (int d = 0; d < 3; ++d) {
                    yb[d] = wi * (t1 * N[d] + t2 * R[d]);
                    Nb[d] = wi * t1 * R[d];
                    gx[d] -= yb[d];
                }
                chain_point(v, xi, eta, w, so, yb, Nb, -wi * S * ri, vb);
            }
        }
        for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) spart[(size_t)p * 12 + 3 * k + d] = vb[k][d];
    }
    for (int d = 0; d < 3; ++d) xs[3 * i + d] = gx[d];
}

// ------------------------------------------------------------------ tangent (JVP)
// Directional derivatives of (y, N, J) at every regular point for vertex velocities dX:
// dy = sum sh_k dv_k, dc = da x b + a x db, dN = o w dc, dJ = w (c . dc) / |c|.
// trec[c][q][8] = {dy (3), dN (3), dJ, 0}.
__global__ void k_tangent_points(int nc, int nq, const double *__restrict__ X,
                                 const double *__restrict__ dX, const int *__restrict__ cell_int,
                                 const int8_t *__restrict__ orient,
                                 const double *__restrict__ qpts, const double *__restrict__ qw,
                                 double *__restrict__ trec)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nc * nq) return;
    const int c = gid / nq, q = gid % nq;
    double sh[4], dx[4], de[4];
    q1_shape(qpts[2 * q], qpts[2 * q + 1], sh, dx, de);
    double a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, da[3] = {0, 0, 0}, db[3] = {0, 0, 0},
           dy[3] = {0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        const int v = cell_int[4 * c + k];
        for (int d = 0; d < 3; ++d) {
            const double p = X[3 * v + d], dp = dX[3 * v + d];
            a[d] += dx[k] * p;
            b[d] += de[k] * p;
            da[d] += dx[k] * dp;
            db[d] += de[k] * dp;
            dy[d] += sh[k] * dp;
        }
    }
    const double cx[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                          a[0] * b[1] - a[1] * b[0]};
    const double dc[3] = {da[1] * b[2] - da[2] * b[1] + a[1] * db[2] - a[2] * db[1],
                          da[2] * b[0] - da[0] * b[2] + a[2] * db[0] - a[0] * db[2],
                          da[0] * b[1] - da[1] * b[0] + a[0] * db[1] - a[1] * db[0]};
    const double cl = sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
    const double so = orient[c] < 0 ? -1.0 : 1.0, w = qw[q];
    double *t = trec + (size_t)gid * 8;
    for (int d = 0; d < 3; ++d) {
        t[d] = dy[d];
        t[3 + d] = so * w * dc[d];
    }
    t[6] = w * (cx[0] * dc[0] + cx[1] * dc[1] + cx[2] * dc[2]) / cl;
    t[7] = 0.0;
}

// part[chunk][row] = sum over the chunk's regular (row, cell) pairs of
// ds/dR . (dy - dx_i) + ds/dN . dN + ds/dJ dJ  (without the 1/(4 pi)); lane = row.
template <int Q>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_jvp_rows(
    const double *__restrict__ xrow, const double *__restrict__ dxrow,
    const int *__restrict__ row_cls, const double *__restrict__ phi,
    const double *__restrict__ cellq, const double *__restrict__ cstate,
    const double *__restrict__ trec, const int4 *__restrict__ cell_cls, int nc,
    int cells_per_chunk, int n, double *__restrict__ part)
{
    constexpr int NQ = Q * Q;
    const int lane = threadIdx.x;
    const int row = blockIdx.x * 64 + lane;
    const int rr = row < n ? row : n - 1;
    const double x0 = xrow[3 * rr], x1 = xrow[3 * rr + 1], x2 = xrow[3 * rr + 2];
    const double e0 = dxrow[3 * rr], e1 = dxrow[3 * rr + 1], e2 = dxrow[3 * rr + 2];
    const int rc = row_cls[rr];
    const double phi_i = phi[rr];
    double acc = 0.0;
    const int c0 = blockIdx.y * cells_per_chunk;
    const int c1 = min(nc, c0 + cells_per_chunk);
    for (int c = c0; c < c1; ++c) {
        const int4 cc = cell_cls[c];
        const double *rec = cellq + (size_t)c * (NQ * 8);
        const double *st = cstate + (size_t)c * (NQ * 2);
        const double *tr = trec + (size_t)c * (NQ * 8);
        double g = 0.0;
        constexpr int B = WB_JVP_BATCH;
#pragma unroll
        for (int q0 = 0; q0 < NQ; q0 += B) {
          double rb[8 * B], tb[8 * B], sb[2 * B];
#pragma unroll
          for (int k = 0; k < 8 * B; ++k) {
              rb[k] = (q0 + k / 8 < NQ) ? rec[8 * q0 + k] : 0.0;
              tb[k] = (q0 + k / 8 < NQ) ? tr[8 * q0 + k] : 0.0;
          }
#pragma unroll
          for (int k = 0; k < 2 * B; ++k) sb[k] = (q0 + k / 2 < NQ) ? st[2 * q0 + k] : 0.0;
#pragma unroll
          for (int u = 0; u < B; ++u) {
            if (q0 + u >= NQ) break;
            const int q = u;  // index into this batch
            const double *r = rb + 8 * u, *t = tb + 8 * u;
            const double R0 = r[0] - x0, R1 = r[1] - x1, R2 = r[2] - x2;
            const double r2 = __builtin_fma(R0, R0, __builtin_fma(R1, R1, R2 * R2));
            const double ri = rsqrt_f64(r2);
            const double ri2 = ri * ri, ri3 = ri2 * ri, ri5 = ri3 * ri2;
            const double RN = __builtin_fma(R0, r[3], __builtin_fma(R1, r[4], R2 * r[5]));
            const double D = sb[2 * q] - phi_i, E = sb[2 * q + 1];
            const double t1 = -D * ri3;
            const double t2 = __builtin_fma(3.0 * D * RN, ri5, E * r[6] * ri3);
            const double d0 = t[0] - e0, d1 = t[1] - e1, d2 = t[2] - e2;  // dR
            const double Nd = __builtin_fma(r[3], d0, __builtin_fma(r[4], d1, r[5] * d2));
            const double Rd = __builtin_fma(R0, d0, __builtin_fma(R1, d1, R2 * d2));
            const double RdN = __builtin_fma(R0, t[3], __builtin_fma(R1, t[4], R2 * t[5]));
            g += t1 * (Nd + RdN) + t2 * Rd - E * ri * t[6];
          }
        }
        const bool sing = (rc == cc.x) | (rc == cc.y) | (rc == cc.z) | (rc == cc.w);
        if (!sing) acc += g;
    }
    if (row < n) part[(size_t)blockIdx.y * n + row] = acc;
}

// Singular pairs of the tangent: ds[i] (without 1/(4 pi)), one thread per row.
__global__ __launch_bounds__(64) void k_jvp_singular(
    const double *__restrict__ X, const double *__restrict__ dX, const int *__restrict__ cell_int,
    const int8_t *__restrict__ orient, const int *__restrict__ sing_ptr,
    const int *__restrict__ sing_cell, const int *__restrict__ sing_vtx, int n, int ns,
    const double *__restrict__ rule, const double *__restrict__ phi,
    const double *__restrict__ psi, double *__restrict__ ds)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi0[3] = {X[3 * i], X[3 * i + 1], X[3 * i + 2]};
    const double dxi[3] = {dX[3 * i], dX[3 * i + 1], dX[3 * i + 2]};
    const double phi_i = phi[i];
    double acc = 0.0;
    for (int p = sing_ptr[i]; p < sing_ptr[i + 1]; ++p) {
        const int c = sing_cell[p], vtx = sing_vtx[p];
        double v[4][3], dv[4][3], pv[4], sv[4];
        for (int k = 0; k < 4; ++k) {
            const int id = cell_int[4 * c + k];
            for (int d = 0; d < 3; ++d) {
                v[k][d] = X[3 * id + d];
                dv[k][d] = dX[3 * id + d];
            }
            pv[k] = phi[id];
            sv[k] = psi[id];
        }
        const double so = orient[c] < 0 ? -1.0 : 1.0;
        const double *ru = rule + (size_t)vtx * ns * 3;
        for (int q = 0; q < ns; ++q) {
            const double xi = ru[3 * q], eta = ru[3 * q + 1], w = ru[3 * q + 2];
            double sh[4], dx[4], de[4];
            q1_shape(xi, eta, sh, dx, de);
            double y[3] = {0, 0, 0}, a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, dy[3] = {0, 0, 0},
                   da[3] = {0, 0, 0}, db[3] = {0, 0, 0};
            double P = 0.0, S = 0.0;
            for (int k = 0; k < 4; ++k) {
                for (int d = 0; d < 3; ++d) {
                    y[d] += sh[k] * v[k][d];
                    a[d] += dx[k] * v[k][d];
                    b[d] += de[k] * v[k][d];
                    dy[d] += sh[k] * dv[k][d];
                    da[d] += dx[k] * dv[k][d];
                    db[d] += de[k] * dv[k][d];
                }
                P += sh[k] * pv[k];
                S += sh[k] * sv[k];
            }
            const double cx[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                                  a[0] * b[1] - a[1] * b[0]};
            const double dc[3] = {da[1] * b[2] - da[2] * b[1] + a[1] * db[2] - a[2] * db[1],
                                  da[2] * b[0] - da[0] * b[2] + a[2] * db[0] - a[0] * db[2],
                                  da[0] * b[1] - da[1] * b[0] + a[0] * db[1] - a[1] * db[0]};
            const double cl = sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
            const double N[3] = {so * w * cx[0], so * w * cx[1], so * w * cx[2]};
            const double dN[3] = {so * w * dc[0], so * w * dc[1], so * w * dc[2]};
            const double J = w * cl;
            const double dJ = w * (cx[0] * dc[0] + cx[1] * dc[1] + cx[2] * dc[2]) / cl;
            const double R[3] = {y[0] - xi0[0], y[1] - xi0[1], y[2] - xi0[2]};
            const double dR[3] = {dy[0] - dxi[0], dy[1] - dxi[1], dy[2] - dxi[2]};
            const double ri = rsqrt_f64(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
            const double ri3 = ri * ri * ri, ri5 = ri3 * ri * ri;
            const double RN = R[0] * N[0] + R[1] * N[1] + R[2] * N[2];
            const double D = P - phi_i;
            const double t1 = -D * ri3, t2 = 3.0 * D * RN * ri5 + S * J * ri3;
            const double Nd = N[0] * dR[0] + N[1] * dR[1] + N[2] * dR[2];
            const double Rd = R[0] * dR[0] + R[1] * dR[1] + R[2] * dR[2];
            const double RdN = R[0] * dN[0] + R[1] * dN[1] + R[2] * dN[2];
            acc += t1 * (Nd + RdN) + t2 * Rd - S * ri * dJ;
        }
    }
    ds[i] = acc;
}

// out[i] = (sum_c part[c][i] + ds[i]) / (4 pi), zero on the rows flagged in skip
__global__ void k_jvp_finish(const double *__restrict__ part, int n_chunks, int n,
                             const double *__restrict__ ds, const uint8_t *__restrict__ skip,
                             double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < n_chunks; ++c) s += part[(size_t)c * n + i];
    out[i] = (skip && skip[i]) ? 0.0 : (s + ds[i]) * INV4PI;
}

// out[e] = sum_c part[c * stride + e]
__global__ void k_sum_chunks(const double *__restrict__ part, int n_chunks, size_t stride,
                             size_t count, double *__restrict__ out)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    double s = 0.0;
    for (int c = 0; c < n_chunks; ++c) s += part[(size_t)c * stride + e];
    out[e] = s;
}

}  // namespace

// 512-row chunks: about 4 x n / 128 wavefronts, enough to keep HBM busy at 10k dofs
int gemvT_chunks(int n) { return max(1, min(128, (n + 511) / 512)); }

void launch_gemvT(int mode, const void *G, const void *H, const void *Glo, const void *Hlo,
                  size_t ld, int n, const uint8_t *gmask, int flip, double s_sel,
                  double s_unsel, const void *x, const double *diag,
                  const uint8_t *constrained, double *part, void *y, hipStream_t s)
{
    const int chunks = gemvT_chunks(n);
    const int rpc = (n + chunks - 1) / chunks;
    dim3 grid((n + GT_COLS - 1) / GT_COLS, chunks);
    const uint16_t *gl = (const uint16_t *)Glo, *hl = (const uint16_t *)Hlo;
    if (mode == 0)
        hipLaunchKernelGGL((k_gemvT<0, double>), grid, dim3(64 * GT_WAVES), 0, s, G, H, gl, hl,
                           ld, n, gmask, flip, s_sel, s_unsel, (const double *)x, constrained,
                           rpc, part);
    else if (mode == 1)
        hipLaunchKernelGGL((k_gemvT<1, float>), grid, dim3(64 * GT_WAVES), 0, s, G, H, gl, hl,
                           ld, n, gmask, flip, s_sel, s_unsel, (const float *)x, constrained,
                           rpc, part);
    else
        hipLaunchKernelGGL((k_gemvT<2, double>), grid, dim3(64 * GT_WAVES), 0, s, G, H, gl, hl,
                           ld, n, gmask, flip, s_sel, s_unsel, (const double *)x, constrained,
                           rpc, part);
    if (mode == 1)
        hipLaunchKernelGGL(k_gemvT_reduce<float>, dim3((n + 255) / 256), dim3(256), 0, s, part,
                           chunks, ld, n, diag, constrained, (const float *)x, (float *)y);
    else
        hipLaunchKernelGGL(k_gemvT_reduce<double>, dim3((n + 255) / 256), dim3(256), 0, s, part,
                           chunks, ld, n, diag, constrained, (const double *)x, (double *)y);
}

void launch_constraint_T(int is_float, int n_targets, const int *tgt, const int *ptr,
                         const int *src, const double *coef, const void *x, void *y,
                         hipStream_t s)
{
    if (n_targets <= 0) return;
    if (is_float)
        hipLaunchKernelGGL(k_constraint_T<float>, dim3((n_targets + 255) / 256), dim3(256), 0, s,
                           n_targets, tgt, ptr, src, coef, (const float *)x, (float *)y);
    else
        hipLaunchKernelGGL(k_constraint_T<double>, dim3((n_targets + 255) / 256), dim3(256), 0,
                           s, n_targets, tgt, ptr, src, coef, (const double *)x, (double *)y);
}

int vjp_row_chunks(int nc) { return max(1, min(64, (nc + 511) / 512)); }
int vjp_point_chunks(int n) { return max(1, min(16, (n + 4095) / 4096)); }

void launch_residual_vjp(int q_order, int n, int nc, const double *X, const int *row_cls,
                         const double *wrow, const double *phi, const double *psi,
                         const double *rowpack, const double *cellq, const int *cell_int,
                         const int *cell_cls, const int8_t *orient, const double *qpts,
                         const double *qw, const int *sing_ptr, const int *sing_cell,
                         const int *sing_vtx, int ns, const double *srule, double *cstate,
                         double *part1, double *part2, double *xgrad, double *vbar,
                         double *xs, double *spart, hipStream_t s)
{
    const int nq = q_order * q_order;
    hipLaunchKernelGGL(k_cell_state, dim3((nc * nq + 255) / 256), dim3(256), 0, s, nc, nq,
                       cell_int, qpts, phi, psi, cstate);
    const int rch = vjp_row_chunks(nc), cpc = (nc + rch - 1) / rch;
    const int pch = vjp_point_chunks(n), rpc = (n + pch - 1) / pch;
    dim3 g1((n + 63) / 64, rch);
    const int cpw = 64 / nq;
    dim3 g2((nc + cpw - 1) / cpw, pch);
#define WB_VJP(QV)                                                                          \
    case QV:                                                                                \
        hipLaunchKernelGGL(k_vjp_rows<QV>, g1, dim3(64), 0, s, X, row_cls, wrow, phi, cellq,  \
                           cstate, (const int4 *)cell_cls, nc, cpc, n, part1);               \
        hipLaunchKernelGGL(k_vjp_points<QV>, g2, dim3(64), 0, s, rowpack, n, rpc, cellq,      \
                           cstate, (const int4 *)cell_cls, nc, part2);                       \
        break;
    switch (q_order) {
        WB_VJP(1)
        WB_VJP(2)
        WB_VJP(3)
        WB_VJP(4)
        WB_VJP(5)
        WB_VJP(6)
    default: break;
    }
#undef WB_VJP
    hipLaunchKernelGGL(k_sum_chunks, dim3((3 * (size_t)n + 255) / 256), dim3(256), 0, s, part1,
                       rch, 3 * (size_t)n, 3 * (size_t)n, xgrad);
    hipLaunchKernelGGL(k_vjp_map, dim3((nc + 127) / 128), dim3(128), 0, s, nc, nq, pch, part2,
                       X, cell_int, orient, qpts, qw, vbar);
    hipLaunchKernelGGL(k_vjp_singular, dim3((n + 63) / 64), dim3(64), 0, s, X, cell_int, orient,
                       sing_ptr, sing_cell, sing_vtx, n, ns, srule, wrow, phi, psi, xs, spart);
}

void launch_residual_jvp(int q_order, int n, int nc, const double *X, const double *dX,
                         const int *row_cls, const double *phi, const double *psi,
                         const double *cellq, const int *cell_int, const int *cell_cls,
                         const int8_t *orient, const double *qpts, const double *qw,
                         const int *sing_ptr, const int *sing_cell, const int *sing_vtx, int ns,
                         const double *srule, const uint8_t *skip, double *cstate,
                         double *trec, double *part, double *ds, double *out, hipStream_t s)
{
    const int nq = q_order * q_order;
    hipLaunchKernelGGL(k_cell_state, dim3((nc * nq + 255) / 256), dim3(256), 0, s, nc, nq,
                       cell_int, qpts, phi, psi, cstate);
    hipLaunchKernelGGL(k_tangent_points, dim3((nc * nq + 255) / 256), dim3(256), 0, s, nc, nq,
                       X, dX, cell_int, orient, qpts, qw, trec);
    const int rch = vjp_row_chunks(nc), cpc = (nc + rch - 1) / rch;
    dim3 g1((n + 63) / 64, rch);
#define WB_JVP(QV)                                                                          \
    case QV:                                                                                \
        hipLaunchKernelGGL(k_jvp_rows<QV>, g1, dim3(64), 0, s, X, dX, row_cls, phi, cellq,    \
                           cstate, trec, (const int4 *)cell_cls, nc, cpc, n, part);          \
        break;
    switch (q_order) {
        WB_JVP(1)
        WB_JVP(2)
        WB_JVP(3)
        WB_JVP(4)
        WB_JVP(5)
        WB_JVP(6)
    default: break;
    }
#undef WB_JVP
    hipLaunchKernelGGL(k_jvp_singular, dim3((n + 63) / 64), dim3(64), 0, s, X, dX, cell_int,
                       orient, sing_ptr, sing_cell, sing_vtx, n, ns, srule, phi, psi, ds);
    hipLaunchKernelGGL(k_jvp_finish, dim3((n + 255) / 256), dim3(256), 0, s, part, rch, n, ds,
                       skip, out);
}

}  // namespace wb
