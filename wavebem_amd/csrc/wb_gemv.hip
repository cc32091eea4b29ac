// wb_gemv.hip -- column-selected GEMV of the BEM operator on gfx950 (HBM-bound).
//
// BEMProblem::vmult (source/bem_problem.cc:621-670) computes
//   dst = -G (surface .* src) + H (other .* src) + alpha .* other .* src
// and compute_rhs (:673-717) the complementary selection.  Every column j needs either
// G(:,j) or H(:,j), never both, so one pass reads exactly one n x n matrix.  The host cuts
// the columns (groups of 4) into segments that read only G, only H, or are mixed (the few
// groups at a Dirichlet/Neumann boundary of the device ordering, resolved per column with a
// 4-bit mask).  A segment is wave-uniform, so the matrix pointer needs no per-lane load and
// all row streams stay in flight.  Each wavefront streams RPW rows with 16-byte lane loads
// (the shapes measured fastest on MI355X: 4 rows x 4 waves for FP64 and FP32, unrolled
// twice) and reduces with wavefront shuffles.  Constrained rows
// (ConstrainedOperator, include/constrained_matrix.h:73-86) are skipped through the
// caller's row list and rewritten by a fix-up kernel.
#include <mutex>
#include <vector>

#include "wb_internal.h"

namespace wb {

namespace {

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef unsigned short uvec4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double bf16d(unsigned short b)
{
    return (double)__uint_as_float((uint32_t)b << 16);
}

// The rows of a wavefront are wave-uniform: keep them (and the row base addresses built
// from them) in SGPRs, so each matrix load is an SGPR base plus the lane's column offset.
template <int RPW>
__device__ __forceinline__ void load_rows(int slot0, const int *__restrict__ rows, int n_rows,
                                          int (&r)[RPW], bool (&ok)[RPW])
{
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int s = __builtin_amdgcn_readfirstlane(slot0 + k);
        ok[k] = s < n_rows;
        r[k] = __builtin_amdgcn_readfirstlane(ok[k] ? (rows ? rows[s] : s) : 0);
    }
}

constexpr int F64_RPW = 4, F64_WAVES = 4;

// segment = int4 {first group, end group, kind (0 G, 1 H, 2 mixed), unused}
__global__ __launch_bounds__(64 * F64_WAVES) void k_gemv_f64(
    const double *__restrict__ G, const double *__restrict__ H, size_t ld,
    const int4 *__restrict__ segs, int n_segs, const uint8_t *__restrict__ gmask, int flip,
    double s_sel, double s_uns, const double *__restrict__ x, const double *__restrict__ diag,
    const double *__restrict__ xd, const int *__restrict__ rows, int n_rows,
    double *__restrict__ y)
{
    constexpr int RPW = F64_RPW;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int r[RPW];
    bool ok[RPW];
    load_rows<RPW>((blockIdx.x * F64_WAVES + wave) * RPW, rows, n_rows, r, ok);
    const double *Gr[RPW], *Hr[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        Gr[k] = G + r[k] * ld;
        Hr[k] = H + r[k] * ld;
    }
    double acc[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) acc[k] = 0.0;
    for (int si = 0; si < n_segs; ++si) {
        const int4 sg = segs[si];
        const int p0 = 2 * sg.x, p1 = 2 * sg.y;  // column pairs of the segment
        if (sg.z < 2) {
            const bool use_h = (sg.z ^ flip) != 0;
            const double sc = use_h ? s_sel : s_uns;
            auto step = [&](int p) {
                const dvec2 xv = *((const dvec2 *)x + p);
#pragma unroll
                for (int k = 0; k < RPW; ++k) {
                    const dvec2 a =
                        __builtin_nontemporal_load((const dvec2 *)(use_h ? Hr[k] : Gr[k]) + p);
                    acc[k] += (a.x * xv.x + a.y * xv.y) * sc;
                }
            };
            int p = p0 + lane;
            for (; p + 64 < p1; p += 128) {
                step(p);
                step(p + 64);
            }
            if (p < p1) step(p);
        } else {
            const unsigned fm = flip ? 3u : 0u;
            for (int p = p0 + lane; p < p1; p += 64) {
                const unsigned m2 = ((gmask[p >> 1] >> ((p & 1) * 2)) & 3u) ^ fm;
                const dvec2 xv = *((const dvec2 *)x + p);
                const double xe0 = xv.x * ((m2 & 1) ? s_sel : s_uns);
                const double xe1 = xv.y * ((m2 & 2) ? s_sel : s_uns);
#pragma unroll
                for (int k = 0; k < RPW; ++k) {
                    const dvec2 a = *((const dvec2 *)Gr[k] + p);
                    const dvec2 b = *((const dvec2 *)Hr[k] + p);
                    acc[k] += ((m2 & 1) ? b.x : a.x) * xe0 + ((m2 & 2) ? b.y : a.y) * xe1;
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const double s = wsum(acc[k]);
        if (lane == 0 && ok[k]) y[r[k]] = s + (diag ? diag[r[k]] * xd[r[k]] : 0.0);
    }
}

// FP32 hi with float vectors (MODE 1) or hi + bf16 lo with double vectors (MODE 2); 4
// columns per lane per step (group g = columns 4g..4g+3).  Both accumulate in FP64: the
// products of FP32 values are exact in FP64 and the kernel is bound by HBM, not by the
// FP64 issue rate, so the FP32 operator's only error is the rounding of its entries.
// Pure segments stream one matrix with vector loads; the few mixed groups load each column
// from its own matrix with scalar loads, so they add no registers to the streaming loop.
template <int MODE> struct MixedShape;
// FP32 shape: 4 rows x 4 waves measured 6.25 TB/s at 41k, 8 x 2 6.22 TB/s
#ifndef WB_GEMV32_RPW
#define WB_GEMV32_RPW 4
#endif
template <> struct MixedShape<1> {
    static constexpr int RPW = WB_GEMV32_RPW, WAVES = 16 / WB_GEMV32_RPW;
};
#ifndef WB_GEMVHL_RPW
#define WB_GEMVHL_RPW 4
#endif
#ifndef WB_GEMVHL_WAVES
#define WB_GEMVHL_WAVES 4
#endif
#ifndef WB_GEMVHL_UNROLL
#define WB_GEMVHL_UNROLL 2
#endif
template <> struct MixedShape<2> {
    static constexpr int RPW = WB_GEMVHL_RPW, WAVES = WB_GEMVHL_WAVES;
};

template <int MODE>
__global__ __launch_bounds__(64 * MixedShape<MODE>::WAVES) void k_gemv_mixed(
    const float *__restrict__ G, const float *__restrict__ H, const uint16_t *__restrict__ Glo,
    const uint16_t *__restrict__ Hlo, size_t ld, const int4 *__restrict__ segs, int n_segs,
    const uint8_t *__restrict__ gmask, int flip, double s_sel, double s_uns,
    const void *__restrict__ xv_, const double *__restrict__ diag, const void *__restrict__ xd_,
    const int *__restrict__ rows, int n_rows, void *__restrict__ y_)
{
    constexpr int RPW = MixedShape<MODE>::RPW, WAVES = MixedShape<MODE>::WAVES;
    typedef double T;  // products and sums in FP64 in both modes
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int r[RPW];
    bool ok[RPW];
    load_rows<RPW>((blockIdx.x * WAVES + wave) * RPW, rows, n_rows, r, ok);
    size_t off[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) off[k] = (size_t)r[k] * ld;
    T acc[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) acc[k] = 0;
    // x of group g, each column scaled by its selection sign
    auto load_x = [&](int g, T (&xe)[4]) {
        if (MODE == 1) {
            const fvec4 xv = *((const fvec4 *)xv_ + g);
            xe[0] = (double)xv.x;
            xe[1] = (double)xv.y;
            xe[2] = (double)xv.z;
            xe[3] = (double)xv.w;
        } else {
            const dvec2 a = *((const dvec2 *)xv_ + 2 * g), b = *((const dvec2 *)xv_ + 2 * g + 1);
            xe[0] = a.x;
            xe[1] = a.y;
            xe[2] = b.x;
            xe[3] = b.y;
        }
    };
    for (int si = 0; si < n_segs; ++si) {
        const int4 sg = segs[si];
        if (sg.z < 2) {
            const bool use_h = (sg.z ^ flip) != 0;
            const T sc = (T)(use_h ? s_sel : s_uns);
            const float *M = use_h ? H : G;
            const uint16_t *Ml = use_h ? Hlo : Glo;
            auto step = [&](int g) {
                T xe[4];
                load_x(g, xe);
#pragma unroll
                for (int e = 0; e < 4; ++e) xe[e] *= sc;
#pragma unroll
                for (int k = 0; k < RPW; ++k) {
                    const fvec4 a = __builtin_nontemporal_load((const fvec4 *)(M + off[k]) + g);
                    if (MODE == 1) {
                        acc[k] += (double)a.x * xe[0] + (double)a.y * xe[1] +
                                  (double)a.z * xe[2] + (double)a.w * xe[3];
                    } else {
                        const uvec4 l =
                            __builtin_nontemporal_load((const uvec4 *)(Ml + off[k]) + g);
                        acc[k] += ((double)a.x + bf16d(l.x)) * xe[0] +
                                  ((double)a.y + bf16d(l.y)) * xe[1] +
                                  ((double)a.z + bf16d(l.z)) * xe[2] +
                                  ((double)a.w + bf16d(l.w)) * xe[3];
                    }
                }
            };
            int g = sg.x + lane;
            constexpr int UN = MODE == 2 ? WB_GEMVHL_UNROLL : 2;  // steps with loads in flight
            for (; g + 64 * (UN - 1) < sg.y; g += 64 * UN) {
#pragma unroll
                for (int u = 0; u < UN; ++u) step(g + 64 * u);
            }
            for (; g < sg.y; g += 64) step(g);
        } else {
            const unsigned fm = flip ? 0xfu : 0u;
            for (int g = sg.x + lane; g < sg.y; g += 64) {
                const unsigned m = gmask[g] ^ fm;
                T xe[4];
                load_x(g, xe);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const bool h = (m >> e) & 1u;
                    const float *M = h ? H : G;
                    const uint16_t *Ml = h ? Hlo : Glo;
                    const T xs = xe[e] * (T)(h ? s_sel : s_uns);
                    const size_t c = (size_t)4 * g + e;
#pragma unroll
                    for (int k = 0; k < RPW; ++k) {
                        T a = (T)M[off[k] + c];
                        if (MODE == 2) a += (T)bf16d(Ml[off[k] + c]);
                        acc[k] += a * xs;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        if (MODE == 1) {
            const double s = wsum(acc[k]);
            if (lane == 0 && ok[k]) {
                const float *xd = (const float *)xd_;
                ((float *)y_)[r[k]] = (float)(s + (diag ? diag[r[k]] * (double)xd[r[k]] : 0.0));
            }
        } else {
            const double s = wsum((double)acc[k]);
            if (lane == 0 && ok[k]) {
                const double *xd = (const double *)xd_;
                ((double *)y_)[r[k]] = s + (diag ? diag[r[k]] * xd[r[k]] : 0.0);
            }
        }
    }
}

// ------------------------------------------------------------------ FP16 inner operator
// A16 = row-scaled FP16 copy of the column-selected vmult operator (H columns of Neumann
// dofs, -G columns of Dirichlet dofs; the alpha diagonal stays separate, as in the FP32
// path), built once per assembly and masks from the FP32 (hi) matrices:
//   A16[i][j] = fp16(A_ij / s_i),  s_i = the power of two at or above max_j |A_ij|,
// so the scaling is exact and every entry lies in [-1, 1].  An inner Krylov solve on A16
// reads 2 bytes per entry; the refinement residuals keep the FP32 hi + bf16 lo operator.
typedef _Float16 hvec8 __attribute__((ext_vector_type(8)));
typedef float fvec8 __attribute__((ext_vector_type(8)));

// FP16 GEMV shape: rows per wavefront x wavefronts per workgroup
#ifndef WB_GEMV16_RPW
#define WB_GEMV16_RPW 4
#endif
#ifndef WB_GEMV16_WAVES
#define WB_GEMV16_WAVES 4
#endif
constexpr int H16_WAVES = WB_GEMV16_WAVES;
constexpr int B16_THREADS = 256;  // one workgroup per row of the build

typedef _Float16 hvec4 __attribute__((ext_vector_type(4)));

// the 4 selected entries of column group g of a row (segment kind: 0 G, 1 H, 2 mixed)
__device__ __forceinline__ fvec4 sel_group(const float *__restrict__ G,
                                           const float *__restrict__ H, size_t off, int g,
                                           int kind, const uint8_t *__restrict__ gmask)
{
    if (kind == 0) return -__builtin_nontemporal_load((const fvec4 *)(G + off) + g);
    if (kind == 1) return __builtin_nontemporal_load((const fvec4 *)(H + off) + g);
    const fvec4 a = *((const fvec4 *)(G + off) + g), b = *((const fvec4 *)(H + off) + g);
    const unsigned m = gmask[g];
    fvec4 r;
    r.x = (m & 1u) ? b.x : -a.x;
    r.y = (m & 2u) ? b.y : -a.y;
    r.z = (m & 4u) ? b.z : -a.z;
    r.w = (m & 8u) ? b.w : -a.w;
    return r;
}

// One workgroup per row: the row maximum over the segments (4 columns per lane per step),
// then the scaled FP16 row.  Columns past n (the last group's tail and the padding up to
// ld) are written as 0.
__global__ __launch_bounds__(B16_THREADS) void k_build_h16(
    const float *__restrict__ G, const float *__restrict__ H, size_t ld, int n,
    const int4 *__restrict__ segs, int n_segs, const uint8_t *__restrict__ gmask,
    _Float16 *__restrict__ A16, float *__restrict__ rs)
{
    __shared__ float red[B16_THREADS / 64];
    const int row = blockIdx.x;
    const size_t off = (size_t)row * ld;
    const int ng = (n + 3) / 4;
    float mx = 0.0f;
    for (int si = 0; si < n_segs; ++si) {
        const int4 sg = segs[si];
        for (int g = sg.x + (int)threadIdx.x; g < sg.y; g += B16_THREADS) {
            const fvec4 v = sel_group(G, H, off, g, sg.z, gmask);
            // entries past n in the last group are zero padding of G and H
            mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    int e = 0;
    if (mx > 0.0f) (void)frexpf(mx, &e);  // mx = f 2^e, f in [0.5, 1): 2^e >= mx
    const float sc = mx > 0.0f ? ldexpf(1.0f, e) : 1.0f, inv = 1.0f / sc;
    for (int si = 0; si < n_segs; ++si) {
        const int4 sg = segs[si];
        for (int g = sg.x + (int)threadIdx.x; g < sg.y; g += B16_THREADS) {
            const fvec4 v = sel_group(G, H, off, g, sg.z, gmask) * inv;
            hvec4 h = __builtin_convertvector(v, hvec4);
            if (4 * g + 3 >= n) {
                if (4 * g + 1 >= n) h.y = (_Float16)0.0f;
                if (4 * g + 2 >= n) h.z = (_Float16)0.0f;
                if (4 * g + 3 >= n) h.w = (_Float16)0.0f;
            }
            *((hvec4 *)(A16 + off) + g) = h;
        }
    }
    for (int g = ng + (int)threadIdx.x; g < (int)(ld / 4); g += B16_THREADS)
        *((hvec4 *)(A16 + off) + g) = hvec4{0, 0, 0, 0};
    if (threadIdx.x == 0) rs[row] = sc;
}

// The same build with the row held in registers: one 1024-thread block per row, thread t
// owns the 4-column groups t + 1024 k (k < K), read once.  The selection comes from the
// group's mask byte (bit e: column e from H), which covers pure and mixed groups alike.
// Same scale and same FP16 values as k_build_h16, at one read of the FP32 entries instead
// of two (the second pass of k_build_h16 misses L2 at these row lengths).
constexpr int BR_THREADS = 1024;
// With S non-null it also writes the row's sums over aggregates of agg4 consecutive
// 4-column groups, S[a ldS + row] = rs_row * sum of the FP16 entries of aggregate a (the
// rows of A P for a coarse space of nc aggregates), from the FP16 values in registers.
template <int K>
__global__ __launch_bounds__(BR_THREADS) void k_build_h16_reg(
    const float *__restrict__ G, const float *__restrict__ H, size_t ld, int n,
    const uint8_t *__restrict__ gmask, _Float16 *__restrict__ A16, float *__restrict__ rs,
    int agg4, int nc, double *__restrict__ S, size_t ldS)
{
    __shared__ float red[BR_THREADS / 64];
    __shared__ float gsum[K * BR_THREADS];
    const int row = blockIdx.x;
    const size_t off = (size_t)row * ld;
    const int ng = (n + 3) / 4;
    fvec4 v[K];
    float mx = 0.0f;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int g = (int)threadIdx.x + BR_THREADS * k;
        v[k] = fvec4{0.f, 0.f, 0.f, 0.f};
        if (g < ng) {
            const unsigned m = gmask[g];
            if (m == 0u) {
                v[k] = -__builtin_nontemporal_load((const fvec4 *)(G + off) + g);
            } else if (m == 15u) {
                v[k] = __builtin_nontemporal_load((const fvec4 *)(H + off) + g);
            } else {
                const fvec4 a = *((const fvec4 *)(G + off) + g), b = *((const fvec4 *)(H + off) + g);
                v[k].x = (m & 1u) ? b.x : -a.x;
                v[k].y = (m & 2u) ? b.y : -a.y;
                v[k].z = (m & 4u) ? b.z : -a.z;
                v[k].w = (m & 8u) ? b.w : -a.w;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
        mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v[k].x), fabsf(v[k].y)),
                             fmaxf(fabsf(v[k].z), fabsf(v[k].w))));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = red[0];
#pragma unroll
    for (int w = 1; w < BR_THREADS / 64; ++w) mx = fmaxf(mx, red[w]);
    int e = 0;
    if (mx > 0.0f) (void)frexpf(mx, &e);  // mx = f 2^e, f in [0.5, 1): 2^e >= mx
    const float sc = mx > 0.0f ? ldexpf(1.0f, e) : 1.0f, inv = 1.0f / sc;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int g = (int)threadIdx.x + BR_THREADS * k;
        if (g < ng) {
            hvec4 h = __builtin_convertvector(v[k] * inv, hvec4);
            if (4 * g + 3 >= n) {
                if (4 * g + 1 >= n) h.y = (_Float16)0.0f;
                if (4 * g + 2 >= n) h.z = (_Float16)0.0f;
                if (4 * g + 3 >= n) h.w = (_Float16)0.0f;
            }
            *((hvec4 *)(A16 + off) + g) = h;
            if (S) gsum[g] = ((float)h.x + (float)h.y) + ((float)h.z + (float)h.w);
        }
    }
    for (int g = ng + (int)threadIdx.x; g < (int)(ld / 4); g += BR_THREADS)
        *((hvec4 *)(A16 + off) + g) = hvec4{0, 0, 0, 0};
    if (threadIdx.x == 0) rs[row] = sc;
    if (!S) return;  // block-uniform
    __syncthreads();
    // all aggregates at once: tpa consecutive threads per aggregate (a power of two with
    // nc tpa <= BR_THREADS), a strided partial sum each, then log2(tpa) shuffle steps
    int tpa = 1;
    while (tpa < 64 && 2 * tpa * nc <= BR_THREADS) tpa *= 2;
    const int a = (int)threadIdx.x / tpa, q = (int)threadIdx.x % tpa;
    double f = 0.0;
    if (a < nc) {
        const int g1 = min(ng, (a + 1) * agg4);
        for (int g = a * agg4 + q; g < g1; g += tpa) f += (double)gsum[g];
    }
    for (int o = 1; o < tpa; o <<= 1) f += __shfl_xor(f, o, 64);
    if (a < nc && q == 0) S[(size_t)a * ldS + row] = f * (double)sc;
}

// y_i = s_i sum_j A16_ij x_j + diag_i xd_i over the listed rows (FP32 vectors, FP32 sums
// per lane, FP64 wavefront reduction); 8 columns per lane per step.
constexpr int H16_RPW = WB_GEMV16_RPW;
__global__ __launch_bounds__(64 * H16_WAVES) void k_gemv_h16(
    const _Float16 *__restrict__ A16, const float *__restrict__ rs, size_t ld, int n8,
    const float *__restrict__ x, const double *__restrict__ diag, const float *__restrict__ xd,
    const int *__restrict__ rows, int n_rows, float *__restrict__ y)
{
    constexpr int RPW = H16_RPW;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int r[RPW];
    bool ok[RPW];
    load_rows<RPW>((blockIdx.x * H16_WAVES + wave) * RPW, rows, n_rows, r, ok);
    size_t off[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) off[k] = (size_t)r[k] * ld;
    float acc[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) acc[k] = 0.0f;
    auto step = [&](int g) {
        const fvec4 x0 = *((const fvec4 *)x + 2 * g), x1 = *((const fvec4 *)x + 2 * g + 1);
#pragma unroll
        for (int k = 0; k < RPW; ++k) {
            const hvec8 a = __builtin_nontemporal_load((const hvec8 *)(A16 + off[k]) + g);
            const fvec8 f = __builtin_convertvector(a, fvec8);
            float t = acc[k];
            t = fmaf(f.s0, x0.x, t);
            t = fmaf(f.s1, x0.y, t);
            t = fmaf(f.s2, x0.z, t);
            t = fmaf(f.s3, x0.w, t);
            t = fmaf(f.s4, x1.x, t);
            t = fmaf(f.s5, x1.y, t);
            t = fmaf(f.s6, x1.z, t);
            t = fmaf(f.s7, x1.w, t);
            acc[k] = t;
        }
    };
    int g = lane;
    for (; g + 64 < n8; g += 128) {
        step(g);
        step(g + 64);
    }
    if (g < n8) step(g);
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const double s = wsum((double)acc[k]);
        if (lane == 0 && ok[k])
            y[r[k]] = (float)(s * (double)rs[r[k]] +
                              (diag ? diag[r[k]] * (double)xd[r[k]] : 0.0));
    }
}

}  // namespace

int resident_blocks(const void *kernel, int threads)
{
    struct Entry {
        const void *k;
        int threads, dev, blocks;
    };
    static std::mutex mu;
    static std::vector<Entry> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lock(mu);
    for (const Entry &e : cache)
        if (e.k == kernel && e.threads == threads && e.dev == dev) return e.blocks;
    int per_cu = 0, cus = 0;
    const bool ok =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess;
    const int blocks = ok && per_cu > 0 && cus > 0 ? per_cu * cus : 0;
    cache.push_back({kernel, threads, dev, blocks});
    return blocks;
}

bool launch_build_h16(const float *G, const float *H, size_t ld, int n, const int *segs,
                      int n_segs, const uint8_t *gmask, void *A16, float *rs, int agg, int nc,
                      double *S, size_t ldS, hipStream_t s)
{
    if (nc > BR_THREADS || agg % 4 != 0) S = nullptr;  // the sums cover one thread per aggregate
    const int agg4 = agg / 4;
    // rows of up to 64k columns are held in registers (read once); longer rows take the
    // two-pass kernel
    const int ng = (n + 3) / 4;
    _Float16 *A = (_Float16 *)A16;
    if (ng <= 4 * BR_THREADS)
        hipLaunchKernelGGL(k_build_h16_reg<4>, dim3(n), dim3(BR_THREADS), 0, s, G, H, ld, n,
                           gmask, A, rs, agg4, nc, S, ldS);
    else if (ng <= 8 * BR_THREADS)
        hipLaunchKernelGGL(k_build_h16_reg<8>, dim3(n), dim3(BR_THREADS), 0, s, G, H, ld, n,
                           gmask, A, rs, agg4, nc, S, ldS);
    else if (ng <= 12 * BR_THREADS)
        hipLaunchKernelGGL(k_build_h16_reg<12>, dim3(n), dim3(BR_THREADS), 0, s, G, H, ld, n,
                           gmask, A, rs, agg4, nc, S, ldS);
    else if (ng <= 16 * BR_THREADS)
        hipLaunchKernelGGL(k_build_h16_reg<16>, dim3(n), dim3(BR_THREADS), 0, s, G, H, ld, n,
                           gmask, A, rs, agg4, nc, S, ldS);
    else {
        hipLaunchKernelGGL(k_build_h16, dim3(n), dim3(B16_THREADS), 0, s, G, H, ld, n,
                           (const int4 *)segs, n_segs, gmask, A, rs);
        return false;  // no aggregate sums
    }
    return S != nullptr;
}

void launch_gemv_h16(const void *A16, const float *rs, size_t ld, int n, const float *x,
                     const double *diag, const float *xd, const int *rows, int n_rows, float *y,
                     hipStream_t s)
{
    if (n_rows <= 0) return;
    constexpr int pb = H16_RPW * H16_WAVES;
    hipLaunchKernelGGL(k_gemv_h16, dim3((n_rows + pb - 1) / pb), dim3(64 * H16_WAVES), 0, s,
                       (const _Float16 *)A16, rs, ld, (n + 7) / 8, x, diag, xd, rows, n_rows, y);
}

void launch_gemv_sel(int mode, const void *G, const void *H, const void *Glo, const void *Hlo,
                     size_t ld, const int *segs, int n_segs, const uint8_t *gmask, int flip,
                     double s_sel, double s_unsel, const void *x, const double *diag,
                     const void *xd, const int *rows, int n_rows, void *y, hipStream_t s)
{
    if (n_rows <= 0) return;
    if (mode == 0) {
        const int per_block = F64_RPW * F64_WAVES;
        hipLaunchKernelGGL(k_gemv_f64, dim3((n_rows + per_block - 1) / per_block),
                           dim3(64 * F64_WAVES), 0, s, (const double *)G, (const double *)H, ld,
                           (const int4 *)segs, n_segs, gmask, flip, s_sel, s_unsel,
                           (const double *)x, diag, (const double *)xd, rows, n_rows, (double *)y);
        return;
    }
    if (mode == 1) {
        constexpr int pb = MixedShape<1>::RPW * MixedShape<1>::WAVES;
        hipLaunchKernelGGL(k_gemv_mixed<1>, dim3((n_rows + pb - 1) / pb),
                           dim3(64 * MixedShape<1>::WAVES), 0, s, (const float *)G,
                           (const float *)H, (const uint16_t *)Glo, (const uint16_t *)Hlo, ld,
                           (const int4 *)segs, n_segs, gmask, flip, s_sel, s_unsel, x, diag, xd,
                           rows, n_rows, y);
    } else {
        constexpr int pb = MixedShape<2>::RPW * MixedShape<2>::WAVES;
        hipLaunchKernelGGL(k_gemv_mixed<2>, dim3((n_rows + pb - 1) / pb),
                           dim3(64 * MixedShape<2>::WAVES), 0, s, (const float *)G,
                           (const float *)H, (const uint16_t *)Glo, (const uint16_t *)Hlo, ld,
                           (const int4 *)segs, n_segs, gmask, flip, s_sel, s_unsel, x, diag, xd,
                           rows, n_rows, y);
    }
}

}  // namespace wb
