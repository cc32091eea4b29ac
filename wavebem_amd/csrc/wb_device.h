// wb_device.h -- device helpers shared by the assembly kernels and their backward.
#pragma once

#include <hip/hip_runtime.h>

namespace wb {

// 1/sqrt(r2) in FP64 from the hardware estimate y0 (relative error eps).
// WB_RSQ_THIRD_ORDER: one step y = y0 (1 + e/2 + 3 e^2/8), e = 1 - r2 y0^2, the series
// of (1 - e)^(-1/2) to second order, leaves about (5/16) eps^3 (5 operations).
// Otherwise two Newton steps y <- y + y (1/2 - (r2/2) y^2), each squaring the error
// (7 operations).
// v_rsq_f64 measured on gfx950: max relative error 5.2e-8 (2^-24.2) over 4M inputs in
// [1e-8, 1e4]; after the third-order step 1.66e-16 (1.5 ulp).
#ifndef WB_RSQ_THIRD_ORDER
#define WB_RSQ_THIRD_ORDER 1
#endif
#ifndef WB_RSQ_NEWTON1
#define WB_RSQ_NEWTON1 0
#endif
__device__ __forceinline__ double rsqrt_f64(double r2)
{
    double y = __builtin_amdgcn_rsq(r2);
#if WB_RSQ_NEWTON1
    // one Newton step: relative error (3/2) eps^2, about 4e-15 (4 operations)
    const double h = 0.5 * r2;
    const double e = __builtin_fma(-h, y * y, 0.5);
    return __builtin_fma(y, e, y);
#elif WB_RSQ_THIRD_ORDER
    const double e = __builtin_fma(-r2, y * y, 1.0);
    const double p = __builtin_fma(0.375, e, 0.5);
    return __builtin_fma(y * e, p, y);
#else
    const double h = 0.5 * r2;
    double t = y * y;
    double e = __builtin_fma(-h, t, 0.5);
    y = __builtin_fma(y, e, y);
    t = y * y;
    e = __builtin_fma(-h, t, 0.5);
    y = __builtin_fma(y, e, y);
    return y;
#endif
}

// 1/sqrt(r2) with one Newton step from the hardware estimate: relative error (3/2) eps^2,
// at most 4.1e-15 with the measured eps above (4 operations).  For values stored as FP32
// hi + bf16 lo (relative rounding 2^-33 = 1.2e-10) this is 30000 times below the storage
// error; the FP64 stores use rsqrt_f64.
__device__ __forceinline__ double rsqrt_f64_newton1(double r2)
{
    const double y = __builtin_amdgcn_rsq(r2);
    const double e = __builtin_fma(-0.5 * r2, y * y, 0.5);
    return __builtin_fma(y, e, y);
}

// Q1 shape values and reference derivatives at (xi, eta), deal.II vertex order
// v0=(0,0) v1=(1,0) v2=(0,1) v3=(1,1).
__device__ __forceinline__ void q1_shape(double xi, double eta, double (&phi)[4],
                                         double (&dxi)[4], double (&deta)[4])
{
    phi[0] = (1 - xi) * (1 - eta);
    phi[1] = xi * (1 - eta);
    phi[2] = (1 - xi) * eta;
    phi[3] = xi * eta;
    dxi[0] = -(1 - eta);
    dxi[1] = (1 - eta);
    dxi[2] = -eta;
    dxi[3] = eta;
    deta[0] = -(1 - xi);
    deta[1] = -xi;
    deta[2] = (1 - xi);
    deta[3] = xi;
}

// In-place inverse of the bs x bs row-major matrix A held in LDS by Gauss-Jordan
// elimination with partial pivoting, by one workgroup of THREADS (>= 64, >= bs) threads.  colk
// (bs doubles), piv (bs ints) and red_i (1 int) are LDS scratch.  Returns, in every
// thread, the number of zero pivots met (0 for a regular matrix).
template <int THREADS, typename T = double>
__device__ int gj_invert_lds(T *A, T *colk, int *piv, int *red_i, int bs)
{
    const int tid = threadIdx.x;
    // elimination layout (bs <= THREADS): thread tid owns column ec = tid % bs of the rows
    // ei, ei + rpp, ... (rpp = THREADS / bs rows per pass); the rest idle there
    const int rpp = THREADS / bs, ec = tid % bs, ei = tid / bs;
    int zero_pivots = 0;
    for (int k = 0; k < bs; ++k) {
        // pivot: argmax |A[i][k]|, i >= k (ties to the smaller row), by the first wavefront
        if (tid < 64) {
            T best = (T)-1;
            int bi = k;
            for (int i = k + tid; i < bs; i += 64) {
                const T a = fabs(A[i * bs + k]);
                if (a > best) {
                    best = a;
                    bi = i;
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const T o = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (o > best || (o == best && oi < bi)) {
                    best = o;
                    bi = oi;
                }
            }
            if (tid == 0) red_i[0] = bi;
        }
        __syncthreads();
        const int p = red_i[0];
        if (tid == 0) piv[k] = p;
        if (p != k)
            for (int c = tid; c < bs; c += THREADS) {
                const T t = A[k * bs + c];
                A[k * bs + c] = A[p * bs + c];
                A[p * bs + c] = t;
            }
        __syncthreads();
        const T d = A[k * bs + k];
        zero_pivots += d == (T)0;
        const T inv = d != (T)0 ? (T)1 / d : (T)0;
        __syncthreads();
        for (int c = tid; c < bs; c += THREADS) A[k * bs + c] = (c == k ? (T)1 : A[k * bs + c]) * inv;
        for (int i = tid; i < bs; i += THREADS) colk[i] = A[i * bs + k];  // before it changes
        __syncthreads();
        if (ei < rpp) {
            const T rk = A[k * bs + ec];  // the scaled pivot row (1/d in column k)
            for (int i = ei; i < bs; i += rpp) {
                if (i == k) continue;
                const T f = colk[i];
                if (ec == k)
                    A[i * bs + ec] = -f * rk;
                else
                    A[i * bs + ec] -= f * rk;
            }
        }
        __syncthreads();
    }
    // undo the row interchanges as column interchanges, last first
    for (int k = bs - 1; k >= 0; --k) {
        const int p = piv[k];
        if (p != k)
            for (int i = tid; i < bs; i += THREADS) {
                const T t = A[i * bs + k];
                A[i * bs + k] = A[i * bs + p];
                A[i * bs + p] = t;
            }
        __syncthreads();
    }
    return zero_pivots;
}

}  // namespace wb
