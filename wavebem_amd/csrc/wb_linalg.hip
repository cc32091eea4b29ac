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
    __shared__ double red_v[BJ_THREADS];
    __shared__ int red_i[BJ_THREADS];
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
    for (int k = 0; k < bs; ++k) {
        // pivot: argmax |A[i][k]|, i >= k
        double best = -1.0;
        int bi = k;
        for (int i = k + tid; i < bs; i += BJ_THREADS) {
            const double a = fabs(A[i * bs + k]);
            if (a > best) {
                best = a;
                bi = i;
            }
        }
        red_v[tid] = best;
        red_i[tid] = bi;
        __syncthreads();
        for (int off = BJ_THREADS / 2; off > 0; off >>= 1) {
            if (tid < off) {
                const double o = red_v[tid + off];
                const int oi = red_i[tid + off];
                if (o > red_v[tid] || (o == red_v[tid] && oi < red_i[tid])) {
                    red_v[tid] = o;
                    red_i[tid] = oi;
                }
            }
            __syncthreads();
        }
        const int p = red_i[0];
        if (tid == 0) piv[k] = p;
        if (p != k)
            for (int c = tid; c < bs; c += BJ_THREADS) {
                const double t = A[k * bs + c];
                A[k * bs + c] = A[p * bs + c];
                A[p * bs + c] = t;
            }
        __syncthreads();
        const double d = A[k * bs + k];
        const double inv = d != 0.0 ? 1.0 / d : 0.0;
        __syncthreads();
        for (int c = tid; c < bs; c += BJ_THREADS) A[k * bs + c] = (c == k ? 1.0 : A[k * bs + c]) * inv;
        for (int i = tid; i < bs; i += BJ_THREADS) colk[i] = A[i * bs + k];  // before it changes
        __syncthreads();
        for (int e = tid; e < bs * bs; e += BJ_THREADS) {
            const int i = e / bs, c = e % bs;
            if (i == k) continue;
            const double f = colk[i];
            if (c == k)
                A[e] = -f * A[k * bs + k];
            else
                A[e] -= f * A[k * bs + c];
        }
        __syncthreads();
    }
    // undo the row interchanges as column interchanges, last first
    for (int k = bs - 1; k >= 0; --k) {
        const int p = piv[k];
        if (p != k)
            for (int i = tid; i < bs; i += BJ_THREADS) {
                const double t = A[i * bs + k];
                A[i * bs + k] = A[i * bs + p];
                A[i * bs + p] = t;
            }
        __syncthreads();
    }
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
template <typename T>
__global__ void k_block_jacobi_apply(int n, int bs, const T *__restrict__ binv,
                                     const T *__restrict__ x, T *__restrict__ y, int k, size_t ldx)
{
    extern __shared__ char smraw[];
    T *xs = (T *)smraw;
    const int b = blockIdx.x, r = threadIdx.x, base = b * bs;
    const int m = min(bs, n - base);
    const T *B = binv + (size_t)b * bs * bs;
    for (int v = 0; v < k; ++v) {
        if (r < bs) xs[r] = (r < m) ? x[(size_t)v * ldx + base + r] : (T)0;
        __syncthreads();
        if (r < m) {
            T acc = 0;
            for (int c = 0; c < m; ++c) acc += B[c * bs + r] * xs[c];
            y[(size_t)v * ldx + base + r] = acc;
        }
        __syncthreads();
    }
}

void launch_block_jacobi_apply(int is_float, int n, int bs, const void *binv, const void *x,
                               void *y, int k, size_t ldx, hipStream_t s)
{
    const int nb = (n + bs - 1) / bs;
    const int threads = bs <= 64 ? 64 : 128;
    if (is_float)
        hipLaunchKernelGGL(k_block_jacobi_apply<float>, dim3(nb), dim3(threads),
                           sizeof(float) * bs, s, n, bs, (const float *)binv, (const float *)x,
                           (float *)y, k, ldx);
    else
        hipLaunchKernelGGL(k_block_jacobi_apply<double>, dim3(nb), dim3(threads),
                           sizeof(double) * bs, s, n, bs, (const double *)binv,
                           (const double *)x, (double *)y, k, ldx);
}

// ------------------------------------------------------------------ Krylov helpers
// Vectors are split into chunks of KCHUNK entries, one workgroup per chunk; reductions
// produce one partial per chunk and are finished, in chunk order, by the consumer, so
// every result is deterministic.
constexpr int KTHREADS = 256;
constexpr int KPER = 4;
constexpr int KCHUNK = KTHREADS * KPER;
constexpr int KMAX = 128;  // max vectors in one orthogonalization

int krylov_partials_blocks(int n) { return (n + KCHUNK - 1) / KCHUNK; }

template <typename T>
__device__ void chunk_dots(const T *__restrict__ V, size_t ldv, int kcount,
                           const T *__restrict__ zs /*LDS chunk*/, int base, int m,
                           double *__restrict__ P, int part)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = wave; k < kcount; k += KTHREADS / 64) {
        const T *vk = V + (size_t)k * ldv + base;
        double s = 0.0;
        for (int e = lane; e < m; e += 64) s += (double)vk[e] * (double)zs[e];
        s = wave_sum(s);
        if (lane == 0) P[(size_t)part * KMAX + k] = s;
    }
}

template <typename T>
__global__ __launch_bounds__(KTHREADS) void k_dot_partials(const T *__restrict__ V, size_t ldv,
                                                           int kcount, const T *__restrict__ z,
                                                           int n, double *__restrict__ P)
{
    __shared__ T zs[KCHUNK];
    const int base = blockIdx.x * KCHUNK;
    const int m = min(KCHUNK, n - base);
    for (int e = threadIdx.x; e < KCHUNK; e += KTHREADS) zs[e] = e < m ? z[base + e] : (T)0;
    __syncthreads();
    chunk_dots(V, ldv, kcount, zs, base, m, P, blockIdx.x);
}

void launch_dot_partials(int is_float, const void *V, size_t ldv, int kcount, const void *z,
                         int n, double *P, hipStream_t s)
{
    dim3 grid(krylov_partials_blocks(n));
    if (is_float)
        hipLaunchKernelGGL(k_dot_partials<float>, grid, dim3(KTHREADS), 0, s, (const float *)V,
                           ldv, kcount, (const float *)z, n, P);
    else
        hipLaunchKernelGGL(k_dot_partials<double>, grid, dim3(KTHREADS), 0, s, (const double *)V,
                           ldv, kcount, (const double *)z, n, P);
}

// h = sum_parts Pin;  z -= V h;  then optional partial dots (Pout) and partial |z|^2
// (stored at Pout[part*KMAX + KMAX-1] when want_norm and Pout has no dots, else in
// hacc[part]).
template <typename T>
__global__ __launch_bounds__(KTHREADS) void k_update_dot(const T *__restrict__ V, size_t ldv,
                                                         int kcount, const double *__restrict__ Pin,
                                                         int nparts, T *__restrict__ z, int n,
                                                         double *__restrict__ Pout,
                                                         double *__restrict__ normp, int want_norm)
{
    __shared__ double h[KMAX];
    __shared__ T zs[KCHUNK];
    __shared__ double red[KTHREADS / 64];
    const int base = blockIdx.x * KCHUNK;
    const int m = min(KCHUNK, n - base);
    for (int k = threadIdx.x; k < kcount; k += KTHREADS) {
        double s = 0.0;
        for (int p = 0; p < nparts; ++p) s += Pin[(size_t)p * KMAX + k];
        h[k] = s;
    }
    __syncthreads();
    double nrm = 0.0;
    for (int e = threadIdx.x; e < KCHUNK; e += KTHREADS) {
        T v = (T)0;
        if (e < m) {
            double a = (double)z[base + e];
            for (int k = 0; k < kcount; ++k) a -= h[k] * (double)V[(size_t)k * ldv + base + e];
            v = (T)a;
            z[base + e] = v;
            nrm += (double)v * (double)v;
        }
        zs[e] = v;
    }
    __syncthreads();
    if (Pout) chunk_dots(V, ldv, kcount, zs, base, m, Pout, blockIdx.x);
    if (want_norm) {
        nrm = wave_sum(nrm);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nrm;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int w = 0; w < KTHREADS / 64; ++w) s += red[w];
            normp[blockIdx.x] = s;
        }
    }
}

void launch_update_dot(int is_float, const void *V, size_t ldv, int kcount, const double *Pin,
                       int nparts, void *z, int n, double *Pout, double *normp, int want_norm,
                       hipStream_t s)
{
    dim3 grid(krylov_partials_blocks(n));
    if (is_float)
        hipLaunchKernelGGL(k_update_dot<float>, grid, dim3(KTHREADS), 0, s, (const float *)V,
                           ldv, kcount, Pin, nparts, (float *)z, n, Pout, normp, want_norm);
    else
        hipLaunchKernelGGL(k_update_dot<double>, grid, dim3(KTHREADS), 0, s, (const double *)V,
                           ldv, kcount, Pin, nparts, (double *)z, n, Pout, normp, want_norm);
}

// hout[k] = sum P1 + sum P2 (k < kcount), hout[kcount] = sqrt(sum Pn); vout = z / hout[kcount]
template <typename T>
__global__ __launch_bounds__(KTHREADS) void k_finalize(int kcount, const double *__restrict__ P1,
                                                       const double *__restrict__ P2,
                                                       const double *__restrict__ Pn, int nparts,
                                                       const T *__restrict__ z, int n,
                                                       T *__restrict__ vout,
                                                       double *__restrict__ hout)
{
    __shared__ double snorm;
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int p = 0; p < nparts; ++p) s += Pn[p];
        snorm = sqrt(s);
    }
    __syncthreads();
    const double nv = snorm;
    if (blockIdx.x == 0) {
        for (int k = threadIdx.x; k < kcount; k += KTHREADS) {
            double s = 0.0;
            for (int p = 0; p < nparts; ++p) s += P1[(size_t)p * KMAX + k];
            if (P2)
                for (int p = 0; p < nparts; ++p) s += P2[(size_t)p * KMAX + k];
            hout[k] = s;
        }
        if (threadIdx.x == 0) hout[kcount] = nv;
    }
    if (vout) {
        const double inv = nv > 0 ? 1.0 / nv : 0.0;
        const int base = blockIdx.x * KCHUNK;
        const int m = min(KCHUNK, n - base);
        for (int e = threadIdx.x; e < m; e += KTHREADS) vout[base + e] = (T)((double)z[base + e] * inv);
    }
}

void launch_finalize(int is_float, int kcount, const double *P1, const double *P2,
                     const double *Pn, int nparts, void *z, int n, void *vout, double *hout,
                     hipStream_t s)
{
    dim3 grid(krylov_partials_blocks(n));
    if (is_float)
        hipLaunchKernelGGL(k_finalize<float>, grid, dim3(KTHREADS), 0, s, kcount, P1, P2, Pn,
                           nparts, (const float *)z, n, (float *)vout, hout);
    else
        hipLaunchKernelGGL(k_finalize<double>, grid, dim3(KTHREADS), 0, s, kcount, P1, P2, Pn,
                           nparts, (const double *)z, n, (double *)vout, hout);
}

// x += sum_i y_i v_i   (x may be double while V is float)
template <typename T, typename TX>
__global__ void k_axpy_basis(const T *__restrict__ V, size_t ldv, int kcount,
                             const double *__restrict__ y, TX *__restrict__ x, int n)
{
    __shared__ double ys[KMAX];
    for (int k = threadIdx.x; k < kcount; k += blockDim.x) ys[k] = y[k];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = (double)x[i];
    for (int k = 0; k < kcount; ++k) a += ys[k] * (double)V[(size_t)k * ldv + i];
    x[i] = (TX)a;
}

void launch_axpy_basis(int is_float, const void *V, size_t ldv, int kcount, const double *y,
                       void *x, int x_is_float, int n, hipStream_t s)
{
    dim3 grid((n + 255) / 256);
    if (is_float && x_is_float)
        hipLaunchKernelGGL((k_axpy_basis<float, float>), grid, dim3(256), 0, s, (const float *)V,
                           ldv, kcount, y, (float *)x, n);
    else if (is_float)
        hipLaunchKernelGGL((k_axpy_basis<float, double>), grid, dim3(256), 0, s, (const float *)V,
                           ldv, kcount, y, (double *)x, n);
    else
        hipLaunchKernelGGL((k_axpy_basis<double, double>), grid, dim3(256), 0, s,
                           (const double *)V, ldv, kcount, y, (double *)x, n);
}

template <typename T>
__global__ __launch_bounds__(KTHREADS) void k_norm2_partials(const T *__restrict__ x, int n,
                                                             double *__restrict__ P)
{
    __shared__ double red[KTHREADS / 64];
    const int base = blockIdx.x * KCHUNK;
    const int m = min(KCHUNK, n - base);
    double s = 0.0;
    for (int e = threadIdx.x; e < m; e += KTHREADS) {
        const double v = (double)x[base + e];
        s += v * v;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < KTHREADS / 64; ++w) t += red[w];
        P[blockIdx.x] = t;
    }
}

void launch_norm2_partials(int is_float, const void *x, int n, double *P, hipStream_t s)
{
    dim3 grid(krylov_partials_blocks(n));
    if (is_float)
        hipLaunchKernelGGL(k_norm2_partials<float>, grid, dim3(KTHREADS), 0, s, (const float *)x,
                           n, P);
    else
        hipLaunchKernelGGL(k_norm2_partials<double>, grid, dim3(KTHREADS), 0, s,
                           (const double *)x, n, P);
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
