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

}  // namespace wb
