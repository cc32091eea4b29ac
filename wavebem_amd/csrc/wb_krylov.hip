// wb_krylov.hip -- Arnoldi building blocks for one or several GMRES instances.
//
// Vectors are split into chunks of KCHUNK entries, one workgroup per chunk; reductions
// produce one partial per chunk and are finished, in chunk order, by the consumer, so every
// result is deterministic.  blockIdx.y selects the instance (right-hand side) of a batch;
// each instance has its own Krylov basis, work vector and partials at the given strides, and
// an optional activity flag lets converged instances sit out.
//
// Orthogonalization is classical Gram-Schmidt with one re-orthogonalization pass (CGS2):
//   h1 = V^T z, z -= V h1, h2 = V^T z, z -= V h2, h = h1 + h2, s = |z|, v_new = z / s
// (dot_partials, then update_dot twice, then finalize); it spans the same Krylov space as
// the modified Gram-Schmidt of deal.II's SolverGMRES with two global reductions per step.
// The chunk partials of a pass are summed by every workgroup of the next kernel, in one
// fixed order (reduce_partials), which saves a launch per reduction; the batched solver
// may instead sum them once with reduce_parts.
#include "wb_internal.h"

namespace wb {

namespace {

constexpr int KTHREADS = 256;
constexpr int KPER = 1;  // 256-entry chunks: about n/256 workgroups per reduction
constexpr int KCHUNK = KTHREADS * KPER;

__device__ __forceinline__ double ksum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

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
        s = ksum(s);
        if (lane == 0) P[(size_t)part * KMAX + k] = s;
    }
}

template <typename T>
__global__ __launch_bounds__(KTHREADS) void k_dot_partials(const T *__restrict__ V, size_t ldv,
                                                           size_t sV, int kcount,
                                                           const T *__restrict__ z, size_t sZ,
                                                           int n, double *__restrict__ P,
                                                           const int *__restrict__ active)
{
    __shared__ T zs[KCHUNK];
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    V += b * sV;
    z += b * sZ;
    P += (size_t)b * gridDim.x * KMAX;
    const int base = blockIdx.x * KCHUNK;
    const int m = min(KCHUNK, n - base);
    for (int e = threadIdx.x; e < KCHUNK; e += KTHREADS) zs[e] = e < m ? z[base + e] : (T)0;
    __syncthreads();
    chunk_dots(V, ldv, kcount, zs, base, m, P, blockIdx.x);
}

// out[k] = sum over the chunks of P[p][k] (k < kcount), one workgroup per instance: thread
// (k, q) sums every KQ-th chunk and the KQ partial sums are added in a fixed order.
constexpr int KQ = KTHREADS / KMAX;
__global__ __launch_bounds__(KTHREADS) void k_reduce_parts(int kcount, const double *__restrict__ P,
                                                           int nparts, double *__restrict__ out,
                                                           size_t sOut,
                                                           const int *__restrict__ active)
{
    __shared__ double red[KQ][KMAX];
    const int b = blockIdx.x;
    if (active && !active[b]) return;
    P += (size_t)b * nparts * KMAX;
    out += b * sOut;
    const int k = threadIdx.x % KMAX, q = threadIdx.x / KMAX;
    double s = 0.0;
    if (k < kcount) {
#pragma unroll 8
        for (int p = q; p < nparts; p += KQ) s += P[(size_t)p * KMAX + k];
    }
    red[q][k] = s;
    __syncthreads();
    if (q == 0 && k < kcount) {
        double t = red[0][k];
        for (int j = 1; j < KQ; ++j) t += red[j][k];
        out[k] = t;
    }
}

// h[k] = sum over the nparts chunk partials P[p][k] (k < kcount), by one workgroup:
// with kw = kcount rounded up to a power of two, thread (k, q) = (t % kw, t / kw) sums
// every (KTHREADS / kw)-th partial from q on, and the KTHREADS / kw sums of each k are then
// added in q order.  The order depends only on kcount and nparts, so every workgroup that
// runs it gets the same bits.
__device__ __forceinline__ void reduce_partials(int kcount, const double *__restrict__ P,
                                                int nparts, double *red /*LDS, KTHREADS*/,
                                                double *h /*LDS, KMAX*/)
{
    int kw = 1;
    while (kw < kcount) kw <<= 1;
    const int nq = KTHREADS / kw;
    const int k = threadIdx.x % kw, q = threadIdx.x / kw;
    double s0 = 0.0, s1 = 0.0;
    if (k < kcount) {
        int p = q;
        for (; p + nq < nparts; p += 2 * nq) {
            s0 += P[(size_t)p * KMAX + k];
            s1 += P[(size_t)(p + nq) * KMAX + k];
        }
        if (p < nparts) s0 += P[(size_t)p * KMAX + k];
    }
    red[threadIdx.x] = s0 + s1;
    __syncthreads();
    if ((int)threadIdx.x < kcount) {
        double t = red[threadIdx.x];
        for (int j = 1; j < nq; ++j) t += red[j * kw + threadIdx.x];
        h[threadIdx.x] = t;
    }
}

// z -= V h;  then optional partial dots (Pout) and partial |z|^2.  h is hin (already
// reduced), or, with Pin, formed here from the nparts chunk partials of the previous pass
// by every workgroup (reduce_partials), workgroup 0 also storing it to hsave.
template <typename T>
__global__ __launch_bounds__(KTHREADS) void k_update_dot(const T *__restrict__ V, size_t ldv,
                                                         size_t sV, int kcount,
                                                         const double *__restrict__ hin,
                                                         size_t sH, T *__restrict__ z,
                                                         size_t sZ, int n,
                                                         double *__restrict__ Pout,
                                                         double *__restrict__ normp,
                                                         int want_norm,
                                                         const int *__restrict__ active,
                                                         const double *__restrict__ Pin,
                                                         double *__restrict__ hsave)
{
    __shared__ double h[KMAX];
    __shared__ T zs[KCHUNK];
    __shared__ double red[KTHREADS / 64];
    __shared__ double rpart[KTHREADS];
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    const int nparts = gridDim.x;
    V += b * sV;
    z += b * sZ;
    if (Pout) Pout += (size_t)b * nparts * KMAX;
    if (normp) normp += (size_t)b * nparts;
    const int base = blockIdx.x * KCHUNK;
    const int m = min(KCHUNK, n - base);
    if (Pin) {
        reduce_partials(kcount, Pin + (size_t)b * nparts * KMAX, nparts, rpart, h);
        __syncthreads();
        if (blockIdx.x == 0)
            for (int k = threadIdx.x; k < kcount; k += KTHREADS) hsave[b * sH + k] = h[k];
    } else {
        hin += b * sH;
        for (int k = threadIdx.x; k < kcount; k += KTHREADS) h[k] = hin[k];
        __syncthreads();
    }
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
        nrm = ksum(nrm);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nrm;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int w = 0; w < KTHREADS / 64; ++w) s += red[w];
            normp[blockIdx.x] = s;
        }
    }
}

// hout[k] = h1[k] + h2[k] (k < kcount), hout[kcount] = sqrt(sum Pn); vout = z / hout[kcount].
// Every workgroup forms the norm itself, wavefront-parallel and in the same order.
// With hmap (coherent pinned host memory, single instance), the column also goes to the
// host: every thread of workgroup 0 writes its entries and fences at system scope, then
// thread 0 publishes seq after them, so a host that reads seq (acquire) sees the column.
// With xcl (single instance), vout is also written to xcl[pos[i]] (FP32).
template <typename T>
__global__ __launch_bounds__(KTHREADS) void k_finalize(int kcount, const double *__restrict__ h1,
                                                       const double *__restrict__ h2,
                                                       size_t sHin,
                                                       const double *__restrict__ Pn,
                                                       const T *__restrict__ z, size_t sZ,
                                                       int n, T *__restrict__ vout, size_t sV,
                                                       double *__restrict__ hout, size_t sH,
                                                       const int *__restrict__ active,
                                                       double *__restrict__ hmap,
                                                       unsigned long long *__restrict__ hseq,
                                                       unsigned long long seq,
                                                       float *__restrict__ xcl,
                                                       const int *__restrict__ pos)
{
    __shared__ double snorm;
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    const int nparts = gridDim.x;
    if (h1) h1 += b * sHin;
    if (h2) h2 += b * sHin;
    Pn += (size_t)b * nparts;
    z += b * sZ;
    if (vout) vout += b * sV;
    hout += b * sH;
    if (threadIdx.x < 64) {
        double s = 0.0;
        for (int p = threadIdx.x; p < nparts; p += 64) s += Pn[p];
        s = ksum(s);
        if (threadIdx.x == 0) snorm = sqrt(s);
    }
    __syncthreads();
    const double nv = snorm;
    if (blockIdx.x == 0) {
        for (int k = threadIdx.x; k < kcount; k += KTHREADS) {
            const double hk = (h1 ? h1[k] : 0.0) + (h2 ? h2[k] : 0.0);
            hout[k] = hk;
            if (hmap) hmap[k] = hk;
        }
        if (threadIdx.x == 0) {
            hout[kcount] = nv;
            if (hmap) hmap[kcount] = nv;
        }
        if (hmap) {
            __threadfence_system();
            __syncthreads();
            if (threadIdx.x == 0)
                __hip_atomic_store(hseq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    if (vout) {
        const double inv = nv > 0 ? 1.0 / nv : 0.0;
        const int base = blockIdx.x * KCHUNK;
        const int m = min(KCHUNK, n - base);
        for (int e = threadIdx.x; e < m; e += KTHREADS) {
            const T v = (T)((double)z[base + e] * inv);
            vout[base + e] = v;
            if (xcl) {
                const int ps = pos[base + e];
                if (ps >= 0) xcl[ps] = (float)v;
            }
        }
    }
}

// x += sum_i y_i v_i   (x may be double while V is float)
template <typename T, typename TX>
__global__ void k_axpy_basis(const T *__restrict__ V, size_t ldv, size_t sV, int kcount,
                             const double *__restrict__ y, size_t sY, TX *__restrict__ x,
                             size_t sX, int n, const int *__restrict__ active)
{
    __shared__ double ys[KMAX];
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    V += b * sV;
    y += b * sY;
    x += b * sX;
    for (int k = threadIdx.x; k < kcount; k += blockDim.x) ys[k] = y[k];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = (double)x[i];
    for (int k = 0; k < kcount; ++k) a += ys[k] * (double)V[(size_t)k * ldv + i];
    x[i] = (TX)a;
}

template <typename T>
__global__ __launch_bounds__(KTHREADS) void k_norm2_partials(const T *__restrict__ x, size_t sX,
                                                             int n, double *__restrict__ P,
                                                             const int *__restrict__ active)
{
    __shared__ double red[KTHREADS / 64];
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    x += b * sX;
    P += (size_t)b * gridDim.x;
    const int base = blockIdx.x * KCHUNK;
    const int m = min(KCHUNK, n - base);
    double s = 0.0;
    for (int e = threadIdx.x; e < m; e += KTHREADS) {
        const double v = (double)x[base + e];
        s += v * v;
    }
    s = ksum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < KTHREADS / 64; ++w) t += red[w];
        P[blockIdx.x] = t;
    }
}

}  // namespace

int krylov_partials_blocks(int n) { return (n + KCHUNK - 1) / KCHUNK; }

void launch_dot_partials(int is_float, const void *V, size_t ldv, int kcount, const void *z,
                         int n, double *P, hipStream_t s, int batch, size_t sV, size_t sZ,
                         const int *active)
{
    dim3 grid(krylov_partials_blocks(n), batch);
    if (is_float)
        hipLaunchKernelGGL(k_dot_partials<float>, grid, dim3(KTHREADS), 0, s, (const float *)V,
                           ldv, sV, kcount, (const float *)z, sZ, n, P, active);
    else
        hipLaunchKernelGGL(k_dot_partials<double>, grid, dim3(KTHREADS), 0, s, (const double *)V,
                           ldv, sV, kcount, (const double *)z, sZ, n, P, active);
}

void launch_reduce_parts(int kcount, const double *P, int n, double *h, hipStream_t s, int batch,
                         size_t sH, const int *active)
{
    if (kcount <= 0) return;
    hipLaunchKernelGGL(k_reduce_parts, dim3(batch), dim3(KTHREADS), 0, s, kcount, P,
                       krylov_partials_blocks(n), h, sH, active);
}

void launch_update_dot(int is_float, const void *V, size_t ldv, int kcount, const double *h,
                       void *z, int n, double *Pout, double *normp, int want_norm, hipStream_t s,
                       int batch, size_t sV, size_t sZ, size_t sH, const int *active,
                       const double *Pin, double *hsave)
{
    dim3 grid(krylov_partials_blocks(n), batch);
    if (is_float)
        hipLaunchKernelGGL(k_update_dot<float>, grid, dim3(KTHREADS), 0, s, (const float *)V,
                           ldv, sV, kcount, h, sH, (float *)z, sZ, n, Pout, normp, want_norm,
                           active, Pin, hsave);
    else
        hipLaunchKernelGGL(k_update_dot<double>, grid, dim3(KTHREADS), 0, s, (const double *)V,
                           ldv, sV, kcount, h, sH, (double *)z, sZ, n, Pout, normp, want_norm,
                           active, Pin, hsave);
}

void launch_finalize(int is_float, int kcount, const double *h1, const double *h2,
                     const double *Pn, void *z, int n, void *vout, double *hout, hipStream_t s,
                     int batch, size_t sZ, size_t sV, size_t sHin, size_t sH,
                     const int *active, double *hmap, unsigned long long *hseq,
                     unsigned long long seq, float *xcl, const int *pos)
{
    dim3 grid(krylov_partials_blocks(n), batch);
    if (is_float)
        hipLaunchKernelGGL(k_finalize<float>, grid, dim3(KTHREADS), 0, s, kcount, h1, h2, sHin,
                           Pn, (const float *)z, sZ, n, (float *)vout, sV, hout, sH, active,
                           hmap, hseq, seq, xcl, pos);
    else
        hipLaunchKernelGGL(k_finalize<double>, grid, dim3(KTHREADS), 0, s, kcount, h1, h2, sHin,
                           Pn, (const double *)z, sZ, n, (double *)vout, sV, hout, sH, active,
                           hmap, hseq, seq, (float *)nullptr, pos);
}

void launch_axpy_basis(int is_float, const void *V, size_t ldv, int kcount, const double *y,
                       void *x, int x_is_float, int n, hipStream_t s, int batch, size_t sV,
                       size_t sY, size_t sX, const int *active)
{
    dim3 grid((n + 255) / 256, batch);
    if (is_float && x_is_float)
        hipLaunchKernelGGL((k_axpy_basis<float, float>), grid, dim3(256), 0, s, (const float *)V,
                           ldv, sV, kcount, y, sY, (float *)x, sX, n, active);
    else if (is_float)
        hipLaunchKernelGGL((k_axpy_basis<float, double>), grid, dim3(256), 0, s,
                           (const float *)V, ldv, sV, kcount, y, sY, (double *)x, sX, n, active);
    else
        hipLaunchKernelGGL((k_axpy_basis<double, double>), grid, dim3(256), 0, s,
                           (const double *)V, ldv, sV, kcount, y, sY, (double *)x, sX, n, active);
}

void launch_norm2_partials(int is_float, const void *x, int n, double *P, hipStream_t s,
                           int batch, size_t sX, const int *active)
{
    dim3 grid(krylov_partials_blocks(n), batch);
    if (is_float)
        hipLaunchKernelGGL(k_norm2_partials<float>, grid, dim3(KTHREADS), 0, s, (const float *)x,
                           sX, n, P, active);
    else
        hipLaunchKernelGGL(k_norm2_partials<double>, grid, dim3(KTHREADS), 0, s,
                           (const double *)x, sX, n, P, active);
}

}  // namespace wb
