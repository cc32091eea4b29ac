// wb_multi.hip -- several right-hand sides at once (MFMA) on gfx950.
//
// Y[:, v] = A_sel X[:, v] (+ diag .* Xd[:, v]) for v < k <= 16 (X, Xd with column stride
// ldx, Y with ldy): the column-selected
// operator of BEMProblem::vmult / compute_rhs (source/bem_problem.cc:621-717) applied to k
// vectors in one pass over the matrix.  Each wavefront owns 16 rows; the product is a dense
// contraction, so it runs on the matrix cores: v_mfma_f32_16x16x4f32 for the FP32 (hi)
// matrix, v_mfma_f64_16x16x4f64 for FP64.  Operand layout (checked on the device):
// A[i][kk] in lane i + 16 kk and B[kk][j] in lane j + 16 kk for both; the result D[i][j]
// is in lane j + 16 (i / 4), register i % 4 for FP32 but in lane j + 16 (i % 4), register
// i / 4 for FP64.  Lane l streams 8 (FP32) or 4 (FP64) consecutive columns of
// row l % 16, so every load is 32 contiguous bytes and a row segment is 128 bytes.
#include "wb_internal.h"

namespace wb {

namespace {

typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef double dvec4 __attribute__((ext_vector_type(4)));

constexpr int MW = 4;  // wavefronts per block, 16 rows each

template <typename T>
__device__ __forceinline__ T row_val(const T *M, size_t off) { return M[off]; }

// FP32 hi matrix, float vectors X (k x ldx), FP32 accumulation on the matrix cores.
__global__ __launch_bounds__(64 * MW) void k_gemm_sel_f32(
    const float *__restrict__ G, const float *__restrict__ H, size_t ld,
    const int4 *__restrict__ segs, int n_segs, const uint8_t *__restrict__ gmask, int flip,
    float s_sel, float s_uns, int k, const float *__restrict__ X, size_t ldx,
    const double *__restrict__ diag, const float *__restrict__ Xd, const int *__restrict__ rows,
    int n_rows, float *__restrict__ Y, size_t ldy)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tile0 = (blockIdx.x * MW + wave) * 16;
    const int i = lane & 15, q = lane >> 4;  // row in tile, k-slot
    const int slot = tile0 + i;
    const int row = slot < n_rows ? (rows ? rows[slot] : slot) : (rows ? rows[0] : 0);
    const int v = lane & 15;  // right-hand side fed by this lane's B operand
    const bool vok = v < k;
    const float *Xv = X + (size_t)(vok ? v : 0) * ldx;
    // The matrix cores accumulate in FP32; every 128 columns the FP32 sums move into FP64
    // accumulators, so rounding grows with the 128-column blocks, not with n.
    fvec4 c = {0.f, 0.f, 0.f, 0.f};
    double cd[4] = {0.0, 0.0, 0.0, 0.0};
    int since = 0;
    for (int si = 0; si < n_segs; ++si) {
        const int4 sg = segs[si];
        const int c_end = 4 * sg.y;
        const unsigned fm = flip ? 0xfu : 0u;
        for (int cb = 4 * sg.x; cb < c_end; cb += 32) {
            if (++since == 4) {
#pragma unroll
                for (int r = 0; r < 4; ++r) cd[r] += (double)c[r];
                c = fvec4{0.f, 0.f, 0.f, 0.f};
                since = 0;
            }
            const int col = cb + 8 * q;  // this lane's 8 columns: col .. col+7
            float a[8], b[8];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int cc = col + 4 * h;
                fvec4 av = {0.f, 0.f, 0.f, 0.f}, xv = {0.f, 0.f, 0.f, 0.f};
                float sgn[4] = {0.f, 0.f, 0.f, 0.f};
                if (cc < c_end) {
                    const int g = cc >> 2;
                    if (sg.z < 2) {
                        const bool use_h = (sg.z ^ flip) != 0;
                        av = __builtin_nontemporal_load(
                            (const fvec4 *)((use_h ? H : G) + (size_t)row * ld) + g);
                        const float s = use_h ? s_sel : s_uns;
                        sgn[0] = sgn[1] = sgn[2] = sgn[3] = s;
                    } else {
                        const unsigned m = gmask[g] ^ fm;
                        const fvec4 ga = *((const fvec4 *)(G + (size_t)row * ld) + g);
                        const fvec4 ha = *((const fvec4 *)(H + (size_t)row * ld) + g);
                        av.x = (m & 1) ? ha.x : ga.x;
                        av.y = (m & 2) ? ha.y : ga.y;
                        av.z = (m & 4) ? ha.z : ga.z;
                        av.w = (m & 8) ? ha.w : ga.w;
#pragma unroll
                        for (int e = 0; e < 4; ++e) sgn[e] = ((m >> e) & 1) ? s_sel : s_uns;
                    }
                    if (vok) xv = *((const fvec4 *)Xv + g);
                }
                a[4 * h + 0] = av.x;
                a[4 * h + 1] = av.y;
                a[4 * h + 2] = av.z;
                a[4 * h + 3] = av.w;
                b[4 * h + 0] = xv.x * sgn[0];
                b[4 * h + 1] = xv.y * sgn[1];
                b[4 * h + 2] = xv.z * sgn[2];
                b[4 * h + 3] = xv.w * sgn[3];
            }
            // step s contracts column col + s of every lane's k-slot q
#pragma unroll
            for (int s = 0; s < 8; ++s) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], b[s], c, 0, 0, 0);
        }
    }
    // D[i][j]: lane holds rows 4q..4q+3 of the tile for right-hand side j = lane & 15
    if (vok) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int s2 = tile0 + 4 * q + r;
            if (s2 < n_rows) {
                const int rr = rows ? rows[s2] : s2;
                double val = cd[r] + (double)c[r];
                if (diag) val += diag[rr] * (double)Xd[(size_t)v * ldx + rr];
                Y[(size_t)v * ldy + rr] = (float)val;
            }
        }
    }
}

// FP64 matrix and vectors on the FP64 matrix cores; lane streams 4 columns (32 bytes).
__global__ __launch_bounds__(64 * MW) void k_gemm_sel_f64(
    const double *__restrict__ G, const double *__restrict__ H, size_t ld,
    const int4 *__restrict__ segs, int n_segs, const uint8_t *__restrict__ gmask, int flip,
    double s_sel, double s_uns, int k, const double *__restrict__ X, size_t ldx,
    const double *__restrict__ diag, const double *__restrict__ Xd,
    const int *__restrict__ rows, int n_rows, double *__restrict__ Y, size_t ldy)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tile0 = (blockIdx.x * MW + wave) * 16;
    const int i = lane & 15, q = lane >> 4;
    const int slot = tile0 + i;
    const int row = slot < n_rows ? (rows ? rows[slot] : slot) : (rows ? rows[0] : 0);
    const int v = lane & 15;
    const bool vok = v < k;
    const double *Xv = X + (size_t)(vok ? v : 0) * ldx;
    dvec4 c = {0.0, 0.0, 0.0, 0.0};
    for (int si = 0; si < n_segs; ++si) {
        const int4 sg = segs[si];
        const int c_end = 4 * sg.y;
        const unsigned fm = flip ? 0xfu : 0u;
        for (int cb = 4 * sg.x; cb < c_end; cb += 16) {
            const int col = cb + 4 * q;  // this lane's 4 columns (one 4-column group)
            double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
            if (col < c_end) {
                const int g = col >> 2;
                double sgn[4];
                if (sg.z < 2) {
                    const bool use_h = (sg.z ^ flip) != 0;
                    const double *M = (use_h ? H : G) + (size_t)row * ld + col;
                    const dvec2 a0 = __builtin_nontemporal_load((const dvec2 *)M);
                    const dvec2 a1 = __builtin_nontemporal_load((const dvec2 *)M + 1);
                    a[0] = a0.x;
                    a[1] = a0.y;
                    a[2] = a1.x;
                    a[3] = a1.y;
                    const double s = use_h ? s_sel : s_uns;
                    sgn[0] = sgn[1] = sgn[2] = sgn[3] = s;
                } else {
                    const unsigned m = gmask[g] ^ fm;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const size_t off = (size_t)row * ld + col + e;
                        a[e] = ((m >> e) & 1) ? H[off] : G[off];
                        sgn[e] = ((m >> e) & 1) ? s_sel : s_uns;
                    }
                }
                if (vok) {
                    const dvec2 x0 = *((const dvec2 *)(Xv + col));
                    const dvec2 x1 = *((const dvec2 *)(Xv + col) + 1);
                    b[0] = x0.x * sgn[0];
                    b[1] = x0.y * sgn[1];
                    b[2] = x1.x * sgn[2];
                    b[3] = x1.y * sgn[3];
                }
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], c, 0, 0, 0);
        }
    }
    // FP64 result layout differs from FP32: D[i][j] in lane j + 16 (i % 4), register i / 4
    if (vok) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int s2 = tile0 + q + 4 * r;
            if (s2 < n_rows) {
                const int rr = rows ? rows[s2] : s2;
                double val = c[r];
                if (diag) val += diag[rr] * Xd[(size_t)v * ldx + rr];
                Y[(size_t)v * ldy + rr] = val;
            }
        }
    }
}

// hi + bf16 lo matrices, FP64 vectors and accumulation, k right-hand sides (VALU; only
// the few refinement residuals use it).  One wavefront per row, lanes over 4-column groups.
__global__ __launch_bounds__(256) void k_gemm_sel_hl(
    const float *__restrict__ G, const float *__restrict__ H, const uint16_t *__restrict__ Glo,
    const uint16_t *__restrict__ Hlo, size_t ld, const int4 *__restrict__ segs, int n_segs,
    const uint8_t *__restrict__ gmask, int flip, double s_sel, double s_uns, int k,
    const double *__restrict__ X, size_t ldx, const double *__restrict__ diag,
    const double *__restrict__ Xd, const int *__restrict__ rows, int n_rows,
    double *__restrict__ Y, size_t ldy)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = blockIdx.x * 4 + wave;
    if (slot >= n_rows) return;
    const int row = rows ? rows[slot] : slot;
    double acc[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0;
    const unsigned fm = flip ? 0xfu : 0u;
    for (int si = 0; si < n_segs; ++si) {
        const int4 sg = segs[si];
        for (int g = sg.x + lane; g < sg.y; g += 64) {
            const unsigned m = (sg.z < 2) ? ((sg.z ^ flip) ? 0xfu : 0u) : (gmask[g] ^ fm);
            for (int e = 0; e < 4; ++e) {
                const size_t off = (size_t)row * ld + 4 * g + e;
                const bool h = (m >> e) & 1;
                const double a = (double)(h ? H[off] : G[off]) +
                                 (double)__uint_as_float((uint32_t)(h ? Hlo[off] : Glo[off]) << 16);
                const double s = h ? s_sel : s_uns;
#pragma unroll
                for (int v = 0; v < 16; ++v)
                    if (v < k) acc[v] += a * s * X[(size_t)v * ldx + 4 * g + e];
            }
        }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        if (v >= k) break;
        double t = acc[v];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        if (lane == 0)
            Y[(size_t)v * ldy + row] = t + (diag ? diag[row] * Xd[(size_t)v * ldx + row] : 0.0);
    }
}

}  // namespace

void launch_gemv_sel_multi(int mode, const void *G, const void *H, const void *Glo,
                           const void *Hlo, size_t ld, const int *segs, int n_segs,
                           const uint8_t *gmask, int flip, double s_sel, double s_unsel, int k,
                           const void *X, size_t ldx, const double *diag, const void *Xd,
                           const int *rows, int n_rows, void *Y, size_t ldy, hipStream_t s)
{
    if (n_rows <= 0 || k <= 0) return;
    if (mode == 2) {
        hipLaunchKernelGGL(k_gemm_sel_hl, dim3((n_rows + 3) / 4), dim3(256), 0, s,
                           (const float *)G, (const float *)H, (const uint16_t *)Glo,
                           (const uint16_t *)Hlo, ld, (const int4 *)segs, n_segs, gmask, flip,
                           s_sel, s_unsel, k, (const double *)X, ldx, diag, (const double *)Xd,
                           rows, n_rows, (double *)Y, ldy);
        return;
    }
    dim3 grid((n_rows + 16 * MW - 1) / (16 * MW)), block(64 * MW);
    if (mode == 1)
        hipLaunchKernelGGL(k_gemm_sel_f32, grid, block, 0, s, (const float *)G, (const float *)H,
                           ld, (const int4 *)segs, n_segs, gmask, flip, (float)s_sel,
                           (float)s_unsel, k, (const float *)X, ldx, diag, (const float *)Xd,
                           rows, n_rows, (float *)Y, ldy);
    else
        hipLaunchKernelGGL(k_gemm_sel_f64, grid, block, 0, s, (const double *)G,
                           (const double *)H, ld, (const int4 *)segs, n_segs, gmask, flip, s_sel,
                           s_unsel, k, (const double *)X, ldx, diag, (const double *)Xd, rows,
                           n_rows, (double *)Y, ldy);
}

}  // namespace wb
