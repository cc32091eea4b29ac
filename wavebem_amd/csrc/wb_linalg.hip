// wb_linalg.hip -- bandwidth-bound kernels of the GMRES path on gfx950.
//
// Column-selected GEMV: BEMProblem::vmult (source/bem_problem.cc:621-670) computes
//   dst = -G (surface .* src) + H (other .* src) + alpha .* other .* src
// and compute_rhs (:673-717) the complementary selection.  Every column j needs either
// G(:,j) or H(:,j), never both, so one kernel reads exactly n entries per row: for each
// group of 4 columns a 4-bit mask says which matrix each column comes from; groups that
// agree (all but the few at a Dirichlet/Neumann boundary in the device ordering) issue
// one vector load from the selected matrix.  Rows run one per (wave, row slot) with
// wavefront shuffle reductions; constrained rows (ConstrainedOperator,
// include/constrained_matrix.h:73-86) are skipped and rewritten by a small fix-up kernel.
#include "wb_internal.h"
#include "wb_device.h"

namespace wb {

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ float wave_sumf(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double bf16_to_double(uint16_t b)
{
    return (double)__uint_as_float((uint32_t)b << 16);
}

// native clang vector types (usable with the nontemporal load builtins)
typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef unsigned short uvec4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ constraint rows
template <typename T>
__global__ void k_constraint_rows(int n_lines, const int *__restrict__ rows,
                                  const int *__restrict__ ptr, const int *__restrict__ cols,
                                  const double *__restrict__ w, const double *__restrict__ inhom,
                                  int mode_inhom, const T *__restrict__ x, T *__restrict__ y)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n_lines) return;
    const int r = rows[l];
    if (mode_inhom) {
        y[r] = (T)inhom[l];
        return;
    }
    double v = (double)x[r];
    for (int p = ptr[l]; p < ptr[l + 1]; ++p) v -= w[p] * (double)x[cols[p]];
    y[r] = (T)v;
}

// op[i] = alpha[i] (1 - s[i]), rhs[i] = -alpha[i] s[i] for a 0/1 mask s
__global__ void k_diag_from_mask(int n, const double *__restrict__ alpha,
                                 const double *__restrict__ s, double *__restrict__ op,
                                 double *__restrict__ rhs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    op[i] = alpha[i] * (1.0 - s[i]);
    rhs[i] = alpha[i] * -s[i];
}

void launch_diag_from_mask(int n, const double *alpha, const double *s, double *op, double *rhs,
                           hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_diag_from_mask, dim3((n + 255) / 256), dim3(256), 0, st, n, alpha, s, op,
                       rhs);
}

// a[i] = v[i] where mask[i] == 0, b[i] = v[i] where mask[i] != 0
__global__ void k_scatter_mask(int n, const uint8_t *__restrict__ mask,
                               const double *__restrict__ v, double *__restrict__ a,
                               double *__restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mask[i]) b[i] = v[i];
    else a[i] = v[i];
}

void launch_scatter_mask(int n, const uint8_t *mask, const double *v, double *a, double *b,
                         hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scatter_mask, dim3((n + 255) / 256), dim3(256), 0, s, n, mask, v, a, b);
}

void launch_constraint_rows(int is_float, int n_lines, const int *rows, const int *ptr,
                            const int *cols, const double *w, const double *inhom, int mode_inhom,
                            const void *x, void *y, hipStream_t s)
{
    if (n_lines <= 0) return;
    dim3 grid((n_lines + 255) / 256);
    if (is_float)
        hipLaunchKernelGGL(k_constraint_rows<float>, grid, dim3(256), 0, s, n_lines, rows, ptr,
                           cols, w, inhom, mode_inhom, (const float *)x, (float *)y);
    else
        hipLaunchKernelGGL(k_constraint_rows<double>, grid, dim3(256), 0, s, n_lines, rows, ptr,
                           cols, w, inhom, mode_inhom, (const double *)x, (double *)y);
}

// ------------------------------------------------------------------ block Jacobi
// Diagonal blocks of the preconditioner matrix of BEMProblem::assemble_preconditioner
// (source/bem_problem.cc:1107-1149): column i holds H(j,i) (+alpha_i on the diagonal)
// for a Neumann column, -G(j,i) for a Dirichlet column; constrained rows are identity
// rows.  Inverted in LDS by Gauss-Jordan elimination with partial pivoting; stored
// column-major (binv[c*bs + r]) so the apply kernel reads it coalesced.
constexpr int BJ_THREADS = 256;

__global__ __launch_bounds__(BJ_THREADS) void k_block_jacobi_build(
    int n, int bs, const void *__restrict__ Gv, const void *__restrict__ Hv,
    const uint16_t *__restrict__ Glo, const uint16_t *__restrict__ Hlo, int precision, size_t ld,
    const double *__restrict__ alpha, const double *__restrict__ surface,
    const uint8_t *__restrict__ constrained, double *__restrict__ binv, float *__restrict__ binv32)
{
    extern __shared__ double sm[];
    double *A = sm;                                // bs x bs, row-major
    double *colk = sm + (size_t)bs * bs;           // bs: column k before elimination
    int *piv = (int *)(sm + (size_t)bs * bs + bs);  // bs
    __shared__ int red_i[1];
    const int b = blockIdx.x;
    const int base = b * bs;
    const int m = min(bs, n - base);
    const int tid = threadIdx.x;
    for (int e = tid; e < bs * bs; e += BJ_THREADS) {
        const int rl = e / bs, cl = e % bs;
        double v = (rl == cl) ? 1.0 : 0.0;
        if (rl < m && cl < m) {
            const int j = base + rl, i = base + cl;
            if (constrained[j]) {
                v = (i == j) ? 1.0 : 0.0;
            } else {
                const size_t off = (size_t)j * ld + i;
                const bool neu = surface[i] == 0.0;
                double gv, hv;
                if (precision == WB_PREC_FP64) {
                    gv = ((const double *)Gv)[off];
                    hv = ((const double *)Hv)[off];
                } else {
                    gv = (double)((const float *)Gv)[off] +
                         (double)__uint_as_float((uint32_t)Glo[off] << 16);
                    hv = (double)((const float *)Hv)[off] +
                         (double)__uint_as_float((uint32_t)Hlo[off] << 16);
                }
                v = neu ? hv + (i == j ? alpha[i] : 0.0) : -gv;
            }
        }
        A[e] = v;
    }
    __syncthreads();
    (void)gj_invert_lds<BJ_THREADS>(A, colk, piv, red_i, bs);
    double *out = binv + (size_t)b * bs * bs;
    float *out32 = binv32 + (size_t)b * bs * bs;
    for (int e = tid; e < bs * bs; e += BJ_THREADS) {
        const int rl = e / bs, cl = e % bs;
        out[cl * bs + rl] = A[e];
        out32[cl * bs + rl] = (float)A[e];
    }
}

void launch_block_jacobi_build(int n, int bs, const void *G, const void *H, const void *Glo,
                               const void *Hlo, int precision, size_t ld, const double *alpha,
                               const double *surface, const uint8_t *constrained, double *binv,
                               float *binv32, hipStream_t s)
{
    const int nb = (n + bs - 1) / bs;
    const size_t smem = sizeof(double) * (bs * bs + bs) + sizeof(int) * bs;
    hipLaunchKernelGGL(k_block_jacobi_build, dim3(nb), dim3(BJ_THREADS), smem, s, n, bs, G, H,
                       (const uint16_t *)Glo, (const uint16_t *)Hlo, precision, ld, alpha,
                       surface, constrained, binv, binv32);
}

// y_blk = Binv_blk x_blk for k vectors (column stride ldx); one thread per block row.
// With C (the coarse-space part of the two-level preconditioner, CS_MAX entries per
// vector): y_i += z_a for the aggregate a = i / agg of row i, z_a = sum_b Aci[a][b] C[v][b]
// (Aci row-major nc x nc; with trans, Aci^T for the transposed system), formed by the
// block for the at most three aggregates its rows touch (bs <= 128, agg >= 64), the
// partial products reduced in a fixed order.
template <typename T>
__global__ void k_block_jacobi_apply(int n, int bs, const T *__restrict__ binv,
                                     const T *__restrict__ x, T *__restrict__ y, int k, size_t ldx,
                                     int trans, const double *__restrict__ C,
                                     const T *__restrict__ Aci, int nc, int agg)
{
    extern __shared__ char smraw[];
    T *xs = (T *)smraw;
    __shared__ double zred[2][3];
    __shared__ double zs[3];
    const int b = blockIdx.x, r = threadIdx.x, base = b * bs;
    const int m = min(bs, n - base);
    const T *B = binv + (size_t)b * bs * bs;
    const int a0 = C ? base / agg : 0, na = C ? (base + m - 1) / agg - a0 + 1 : 0;
    for (int v = 0; v < k; ++v) {
        if (r < bs) xs[r] = (r < m) ? x[(size_t)v * ldx + base + r] : (T)0;
        for (int q = 0; q < na; ++q) {
            const int a = a0 + q;
            const double *c = C + (size_t)v * CS_MAX;
            double p = 0.0;
            // row a of Ac^-1, or with trans (the adjoint) column a, i.e. row a of Ac^-T
            for (int bb = r; bb < nc; bb += blockDim.x)
                p += (double)(trans ? Aci[(size_t)bb * nc + a] : Aci[(size_t)a * nc + bb]) * c[bb];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off, 64);
            if ((r & 63) == 0) zred[r >> 6][q] = p;
        }
        __syncthreads();
        if (r < na) zs[r] = blockDim.x > 64 ? zred[0][r] + zred[1][r] : zred[0][r];
        __syncthreads();
        if (r < m) {
            T acc = 0;
            if (trans)  // the transposed block inverse, for the adjoint system
                for (int c = 0; c < m; ++c) acc += B[r * bs + c] * xs[c];
            else
                for (int c = 0; c < m; ++c) acc += B[c * bs + r] * xs[c];
            if (C) acc = (T)((double)acc + zs[(base + r) / agg - a0]);
            y[(size_t)v * ldx + base + r] = acc;
        }
        __syncthreads();
    }
}

void launch_block_jacobi_apply(int is_float, int n, int bs, const void *binv, const void *x,
                               void *y, int k, size_t ldx, hipStream_t s, int trans,
                               const double *C, const void *Aci, int nc, int agg)
{
    const int nb = (n + bs - 1) / bs;
    const int threads = bs <= 64 ? 64 : 128;
    if (is_float)
        hipLaunchKernelGGL(k_block_jacobi_apply<float>, dim3(nb), dim3(threads),
                           sizeof(float) * bs, s, n, bs, (const float *)binv, (const float *)x,
                           (float *)y, k, ldx, trans, C, (const float *)Aci, nc, agg);
    else
        hipLaunchKernelGGL(k_block_jacobi_apply<double>, dim3(nb), dim3(threads),
                           sizeof(double) * bs, s, n, bs, (const double *)binv,
                           (const double *)x, (double *)y, k, ldx, trans, C,
                           (const double *)Aci, nc, agg);
}

// elementwise ops on n entries:
//  0: c(d) = a(d) - b(d)          1: c(f) = (float) a(d)          2: c(d) += a(f)
//  3: c(d) = a(d) * alpha         4: c(d) = a(d) * b(d)            5: c(f) = a(f) * alpha
//  6: c(d) = 0                    7: c(f) = 0                      8: c(d) = b(d) - a(d)
//  9: c(f) = b(f) - a(f)
__global__ void k_vec_ops(int op, int n, const void *__restrict__ a, const void *__restrict__ b,
                          void *__restrict__ c, double alpha)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *ad = (const double *)a, *bd = (const double *)b;
    const float *af = (const float *)a;
    double *cd = (double *)c;
    float *cf = (float *)c;
    switch (op) {
    case 0: cd[i] = ad[i] - bd[i]; break;
    case 1: cf[i] = (float)ad[i]; break;
    case 2: cd[i] += (double)af[i]; break;
    case 3: cd[i] = ad[i] * alpha; break;
    case 4: cd[i] = ad[i] * bd[i]; break;
    case 5: cf[i] = (float)((double)af[i] * alpha); break;
    case 6: cd[i] = 0.0; break;
    case 7: cf[i] = 0.0f; break;
    case 8: cd[i] = bd[i] - ad[i]; break;
    case 9: cf[i] = ((const float *)b)[i] - af[i]; break;
    default: break;
    }
}

void launch_vec_ops(int op, int n, const void *a, const void *b, void *c, double alpha,
                    hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_vec_ops, dim3((n + 255) / 256), dim3(256), 0, s, op, n, a, b, c, alpha);
}

// gather: dst[i] = src[idx[i]];  scatter: dst[idx[i]] = src[i]
__global__ void k_permute(int n, const int *__restrict__ idx, const double *__restrict__ src,
                          double *__restrict__ dst, int gather)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (gather)
        dst[i] = src[idx[i]];
    else
        dst[idx[i]] = src[i];
}

void launch_permute(int n, const int *idx, const double *src, double *dst, int gather,
                    hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_permute, dim3((n + 255) / 256), dim3(256), 0, s, n, idx, src, dst, gather);
}

}  // namespace wb
