// wb_hmat.hip -- compressed copy of the vmult operator for the inner solves of FP32
// iterative refinement, and its product (gfx950).
//
// The refinement residual is always formed with the FP64-accurate operator (hi + lo), so
// the inner correction solve only has to reduce the residual by a few orders per sweep;
// an operator accurate to ~1e-4 of each block's largest entry serves it as well as the
// row-scaled FP16 copy, at a fraction of the bytes per product.  The rows and columns are
// split into clusters of <= HM_CL dofs (plan_hmat).  Block (I, J) is stored
//   * dense, FP16 scaled by a power of two >= its largest entry, when the clusters touch or
//     the low-rank test fails;
//   * as Q B, Q (|I| x r) with orthonormal columns and B = Q^T A_IJ (r x |J|), FP16, with
//     r <= HM_RMAX chosen by a randomized range finder: Y = A_IJ S with HM_SK random sign
//     vectors, then a column-pivoted Cholesky of the Gram matrix Y^T Y (FP64, on the matrix
//     cores) until the largest residual column of Y (Halko, Martinsson & Tropp 2011,
//     section 4.3) is below tol * max|A_IJ| * sqrt(|I| |J|), i.e. an entry RMS error of
//     about tol times the block's largest entry;
//   * not at all when it is zero (rank 0).
// Build per assembly (and mask change): k_hm_sketch (one pass over the FP32 hi rows),
// k_hm_qr (one wavefront per block), k_hm_fill (a second pass over the rows of each block).
// Product: k_hm_gemv (one wavefront per 64 rows of a row cluster and a chunk of column
// clusters) and k_hm_finish (chunk sums, the alpha diagonal).
#include <algorithm>

#include "wb_internal.h"

namespace wb {

namespace {

typedef _Float16 hvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef double dvec4 __attribute__((ext_vector_type(4)));

// Wavefront sum / max with DPP (quad permutes, row mirrors, row broadcasts): the total
// lands in lane 63.
template <bool MAX>
__device__ __forceinline__ float dpp_step(float v, float t)
{
    return MAX ? fmaxf(v, t) : v + t;
}

#define WB_DPP_MOV(x, ctrl) \
    __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), ctrl, 0xf, 0xf, false))
#define WB_DPP_ROWS(x, ctrl, rows) \
    __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), ctrl, rows, 0xf, false))

template <bool MAX>
__device__ __forceinline__ float wave_reduce63(float v)
{
    v = dpp_step<MAX>(v, WB_DPP_MOV(v, 0xb1));   // quad_perm [1,0,3,2]
    v = dpp_step<MAX>(v, WB_DPP_MOV(v, 0x4e));   // quad_perm [2,3,0,1]
    v = dpp_step<MAX>(v, WB_DPP_MOV(v, 0x141));  // row_half_mirror
    v = dpp_step<MAX>(v, WB_DPP_MOV(v, 0x140));  // row_mirror: every lane holds its row's total
    // rows 1 and 3 take the last lane of the row before (row_bcast:15), then rows 2 and 3
    // the last lane of row 1 (row_bcast:31); the other rows read 0 and keep their value
    const float b15 = WB_DPP_ROWS(v, 0x142, 0xa);
    v = MAX ? ((threadIdx.x & 16u) ? fmaxf(v, b15) : v) : v + b15;
    const float b31 = WB_DPP_ROWS(v, 0x143, 0xc);
    v = MAX ? ((threadIdx.x & 32u) ? fmaxf(v, b31) : v) : v + b31;
    return v;
}

__device__ __forceinline__ float lane63(float v)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// The vmult operator's FP32 hi entry (row offset off, device column col): H for a Neumann
// column (gmask bit set), -G for a Dirichlet column.
__device__ __forceinline__ float sel_entry(const float *__restrict__ G,
                                           const float *__restrict__ H,
                                           const uint8_t *__restrict__ gmask, size_t off, int col)
{
    const bool h = (gmask[col >> 2] >> (col & 3)) & 1u;
    return h ? H[off + col] : -G[off + col];
}

// sign of sketch vector k at cluster-local position p: bit k of a 32-bit integer hash of p
__device__ __forceinline__ uint32_t sign_bits(uint32_t p)
{
    uint32_t h = p * 0x9e3779b1u + 0x7f4a7c15u;
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    h ^= h >> 12;
    h *= 0x297a2d39u;
    h ^= h >> 15;
    return h;
}

__device__ __forceinline__ float pow2_at_least(float m)
{
    int e = 0;
    if (!(m > 0.0f)) return 1.0f;
    (void)frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
    return ldexpf(1.0f, e);
}

}  // namespace

// One wavefront per 16 consecutive device rows, all column clusters J in order, on the
// matrix cores (v_mfma_f32_16x16x4f32: A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k,
// D[i][j] in lane j + 16 (i / 4), register i % 4):
//   Y[J][row][k] = sum over the columns p of cluster J of A(row, col_J(p)) s_k(p),
//   Mx[J][row]   = max |A(row, col_J(p))|,
// lane (i, kk) reading row i at the positions q = 16 m + 4 kk + e of each segment.  The
// segment descriptors {first column, length, kind, 0} say whether the segment reads only G,
// only H or both (then column by column through gmask), so the loads of a whole cluster
// depend on nothing and are issued together.  With S, also S[a][row] = the row's sum over
// the columns of aggregate a (agg consecutive device columns, agg >= HM_SEG, so a segment
// meets at most two aggregates), summed per segment in a fixed order.
__global__ __launch_bounds__(64) void k_hm_sketch(const float *__restrict__ G,
                                                  const float *__restrict__ H, size_t ld, int n,
                                                  const uint8_t *__restrict__ gmask,
                                                  const int4 *__restrict__ cl_seg, int ncl,
                                                  float *__restrict__ Y, float *__restrict__ Mx,
                                                  int agg, int nagg, double *__restrict__ S,
                                                  size_t ldS)
{
    __shared__ float sacc[16][CS_MAX];
    const int lane = threadIdx.x, i = lane & 15, kk = lane >> 4;
    const int row0 = blockIdx.x * 16, row = row0 + i;
    const bool rok = row < n;
    const size_t off = (size_t)(rok ? row : 0) * ld;
    // the lane's sign of sketch vector i at position 16 g + 4 kk + e: bit 4 g + e of smask
    uint64_t smask = 0;
#pragma unroll
    for (int g = 0; g < 16; ++g)
#pragma unroll
        for (int e = 0; e < 4; ++e)
            smask |= (uint64_t)((sign_bits((uint32_t)(16 * g + 4 * kk + e)) >> i) & 1u)
                     << (4 * g + e);
    uint32_t sm0 = (uint32_t)smask, sm1 = (uint32_t)(smask >> 32);
    auto sgn_of = [&](int bit) {
        const uint32_t w = bit < 32 ? sm0 : sm1;
        return __uint_as_float(0xbf800000u ^ (((w >> (bit & 31)) & 1u) << 31));  // bit ? 1 : -1
    };
    if (S)
        for (int t = lane; t < 16 * CS_MAX; t += 64) sacc[t / CS_MAX][t % CS_MAX] = 0.0f;
    __syncthreads();
    for (int J = 0; J < ncl; ++J) {
        // keep the 64 signs out of the registers: recomputed per cluster from the mask
        asm volatile("" : "+v"(sm0), "+v"(sm1));
        int4 sd[HM_SEGS];
#pragma unroll
        for (int q = 0; q < HM_SEGS; ++q) {
            sd[q] = cl_seg[J * HM_SEGS + q];
            sd[q].x = __builtin_amdgcn_readfirstlane(sd[q].x);
            sd[q].y = __builtin_amdgcn_readfirstlane(sd[q].y);
            sd[q].z = __builtin_amdgcn_readfirstlane(sd[q].z);
        }
        // all loads of the cluster first: lane (i, kk) reads 16 contiguous bytes of row i
        // at segment positions 16 m + 4 kk .. + 3 (dword-aligned 16-byte loads); a tail
        // past the segment's length is read element by element, so no load leaves the row
        float a[HM_SEGS][16];
#pragma unroll
        for (int q = 0; q < HM_SEGS; ++q) {
            const int start = sd[q].x, len = sd[q].y, kind = sd[q].z;
            const float *src = (kind == 1 ? H : G) + off + start;
            const float sgn = kind == 1 ? 1.0f : -1.0f;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int p = 16 * m + 4 * kk;
                fvec4 v = {0.f, 0.f, 0.f, 0.f};
                if (rok && p < len) {
                    if (kind != 2) {
                        if (p + 4 <= len) {
                            __builtin_memcpy(&v, src + p, 16);
                        } else {
#pragma unroll
                            for (int e = 0; e < 4; ++e)
                                if (p + e < len) v[e] = src[p + e];
                        }
                        v *= sgn;
                    } else {
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            if (p + e < len) {
                                const int col = start + p + e;
                                const bool h = (gmask[col >> 2] >> (col & 3)) & 1u;
                                const float w = (h ? H : G)[off + col];
                                v[e] = h ? w : -w;
                            }
                    }
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) a[q][4 * m + e] = v[e];
            }
        }
        fvec4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
        float mx = 0.0f;
#pragma unroll
        for (int q = 0; q < HM_SEGS; ++q) {
            float slo = 0.0f, shi = 0.0f;
            const int abound = S ? (sd[q].x / agg + 1) * agg : 0;
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const float v = a[q][t];
                mx = fmaxf(mx, fabsf(v));
                if (S) {
                    const int col = sd[q].x + 16 * (t >> 2) + 4 * kk + (t & 3);
                    if (col < abound)
                        slo += v;
                    else
                        shi += v;
                }
                const float sb = sgn_of(4 * (4 * q + (t >> 2)) + (t & 3));
                if (t & 1)
                    acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(v, sb, acc1, 0, 0, 0);
                else
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(v, sb, acc0, 0, 0, 0);
            }
            if (S && sd[q].y > 0) {
                slo += __shfl_xor(slo, 16, 64);
                slo += __shfl_xor(slo, 32, 64);
                shi += __shfl_xor(shi, 16, 64);
                shi += __shfl_xor(shi, 32, 64);
                if (kk == 0) {
                    const int a0 = sd[q].x / agg;
                    sacc[i][a0] += slo;
                    if (a0 + 1 < nagg) sacc[i][a0 + 1] += shi;
                }
            }
        }
        const fvec4 acc = acc0 + acc1;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int r = row0 + 4 * kk + g;
            if (r < n) Y[((size_t)J * n + r) * HM_SK + i] = acc[g];
        }
        mx = fmaxf(mx, __shfl_xor(mx, 16, 64));
        mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
        if (kk == 0 && rok) Mx[(size_t)J * n + row] = mx;
    }
    if (S) {
        __syncthreads();
        for (int t = lane; t < 16 * nagg; t += 64) {
            const int r = row0 + t % 16, a = t / 16;
            if (r < n) S[(size_t)a * ldS + r] = (double)sacc[t % 16][a];
        }
    }
}

// One wavefront per block (I, J): the block's largest entry (fscale = a power of two >= it),
// then for blocks not dense by geometry the rank r and Q (HM_RMAX x HM_CL FP16, column k
// of Q at row k) from the sketch rows of I.
__global__ __launch_bounds__(64) void k_hm_qr(const float *__restrict__ Y,
                                              const float *__restrict__ Mx, int n,
                                              const int *__restrict__ cl_idx,
                                              const int *__restrict__ cl_size, int ncl,
                                              const uint8_t *__restrict__ near, float tol,
                                              uint8_t *__restrict__ rank,
                                              float *__restrict__ fscale,
                                              _Float16 *__restrict__ Q16)
{
    __shared__ double Gm[HM_SK][HM_SK + 1];
    __shared__ double Rm[HM_RMAX][HM_SK];
    __shared__ int piv[HM_RMAX];
    const int b = blockIdx.x, lane = threadIdx.x;
    const int I = b / ncl, J = b % ncl;
    const int szI = cl_size[I], szJ = cl_size[J];
    const int *rows = cl_idx + (size_t)I * HM_CL;
    float m = 0.0f;
    for (int i = lane; i < szI; i += 64) m = fmaxf(m, Mx[(size_t)J * n + rows[i]]);
    m = lane63(wave_reduce63<true>(m));
    if (lane == 0) fscale[b] = pow2_at_least(m);
    if (near[b]) {
        if (lane == 0) rank[b] = (uint8_t)HM_DENSE;
        return;
    }
    if (!(m > 0.0f)) {
        if (lane == 0) rank[b] = 0;
        return;
    }
    // Gram matrix Y^T Y on the matrix cores: A[a][kk] = B[kk][a] = Y(row 4t + kk, a), held
    // by lane a + 16 kk for both operands; the result D[i][j] is in lane j + 16 (i % 4),
    // register i / 4
    dvec4 acc = {0.0, 0.0, 0.0, 0.0};
    const int a = lane & 15, kk = lane >> 4;
    for (int t = 0; 4 * t < szI; ++t) {
        const int li = 4 * t + kk;
        const double v = li < szI ? (double)Y[((size_t)J * n + rows[li]) * HM_SK + a] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v, v, acc, 0, 0, 0);
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) Gm[4 * g + kk][a] = acc[g];
    __syncthreads();
    // column-pivoted Cholesky: lane j < HM_SK owns column j
    const double thresh2 = (double)tol * (double)tol * (double)m * (double)m * (double)szI *
                           (double)szJ;
    double d = lane < HM_SK ? Gm[lane][lane] : -1.0;
    bool chosen = lane >= HM_SK;
    int r = 0;
    for (;;) {
        // argmax of d over the unchosen columns (ties: the lower index)
        double best = chosen ? -1.0 : d;
        int bi = lane;
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o, 16);
            const int oi = __shfl_xor(bi, o, 16);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        best = __shfl(best, 0, 64);
        bi = __shfl(bi, 0, 64);
        if (!(best > thresh2)) break;
        if (r == HM_RMAX) {
            r = HM_DENSE;
            break;
        }
        const double rpp = sqrt(best);
        if (lane < HM_SK) {
            double v;
            if (lane == bi) {
                v = rpp;
                chosen = true;
            } else {
                double s = Gm[bi][lane];
                for (int q = 0; q < r; ++q) s -= Rm[q][bi] * Rm[q][lane];
                v = chosen ? 0.0 : s / rpp;
                if (!chosen) d -= v * v;
            }
            Rm[r][lane] = v;
        }
        if (lane == 0) piv[r] = bi;
        __syncthreads();
        ++r;
    }
    if (lane == 0) rank[b] = (uint8_t)r;
    if (r == HM_DENSE || r == 0) return;
    // Q = Y(:, piv) R^-1, row by row
    for (int i = lane; i < szI; i += 64) {
        const float *y = Y + ((size_t)J * n + rows[i]) * HM_SK;
        double q[HM_RMAX];
#pragma unroll
        for (int k = 0; k < HM_RMAX; ++k) {
            q[k] = 0.0;
            if (k < r) {
                const int pk = piv[k];
                double s = (double)y[pk];
#pragma unroll
                for (int m2 = 0; m2 < HM_RMAX; ++m2)
                    if (m2 < k) s -= Rm[m2][pk] * q[m2];
                q[k] = s / Rm[k][pk];
                Q16[((size_t)b * HM_RMAX + k) * HM_CL + i] = (_Float16)q[k];
            }
        }
    }
}

// One wavefront per block: dense blocks -> D16 (rows of the block, |J| columns, scaled by
// 1 / fscale); low-rank blocks -> B16 = (Q16^T A_IJ) / fscale, fscale replaced by a power
// of two >= max |B|.  Lane l holds the block columns 4 l .. 4 l + 3 (their matrix and sign
// chosen once) and the block rows l, l + 64, ... (read back with readlane), so the loads
// of a row depend on nothing but the row.
__global__ __launch_bounds__(64) void k_hm_fill(const float *__restrict__ G,
                                                const float *__restrict__ H, size_t ld,
                                                const uint8_t *__restrict__ gmask,
                                                const int *__restrict__ cl_idx,
                                                const int *__restrict__ cl_size, int ncl,
                                                const uint8_t *__restrict__ rank,
                                                float *__restrict__ fscale,
                                                const _Float16 *__restrict__ Q16,
                                                _Float16 *__restrict__ B16,
                                                _Float16 *__restrict__ D16)
{
    __shared__ float Qs[HM_RMAX][HM_CL];
    const int b = blockIdx.x, lane = threadIdx.x;
    const int I = b / ncl, J = b % ncl;
    const int r = rank[b];
    if (r == 0) return;
    const int szI = cl_size[I], szJ = cl_size[J];
    int rowv[HM_CL / 64];
#pragma unroll
    for (int u = 0; u < HM_CL / 64; ++u) {
        const int li = lane + 64 * u;
        rowv[u] = li < szI ? cl_idx[(size_t)I * HM_CL + li] : 0;
    }
    const float *ptr[4];
    float sgn[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int p = 4 * lane + c;
        if (p < szJ) {
            const int col = cl_idx[(size_t)J * HM_CL + p];
            const bool h = (gmask[col >> 2] >> (col & 3)) & 1u;
            ptr[c] = (h ? H : G) + col;
            sgn[c] = h ? 1.0f : -1.0f;
        } else {
            ptr[c] = G;
            sgn[c] = 0.0f;
        }
    }
    auto row_off = [&](int i) {
        int v = rowv[0];
#pragma unroll
        for (int u = 1; u < HM_CL / 64; ++u)
            if ((i >> 6) == u) v = rowv[u];
        return (size_t)__builtin_amdgcn_readlane(v, i & 63) * ld;
    };
    if (r == HM_DENSE) {
        const float inv = 1.0f / fscale[b];
        for (int i0 = 0; i0 < szI; i0 += 8) {
            float v[8][4];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = min(i0 + u, szI - 1);
                const size_t off = row_off(i);
#pragma unroll
                for (int c = 0; c < 4; ++c) v[u][c] = ptr[c][off] * sgn[c];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (i0 + u < szI) {
                    hvec4 h;
#pragma unroll
                    for (int c = 0; c < 4; ++c) h[c] = (_Float16)(v[u][c] * inv);
                    *(hvec4 *)(D16 + ((size_t)b * HM_CL + i0 + u) * HM_CL + 4 * lane) = h;
                }
        }
        return;
    }
    for (int e = lane; e < r * HM_CL; e += 64) {
        const int k = e / HM_CL, i = e % HM_CL;
        Qs[k][i] = i < szI ? (float)Q16[((size_t)b * HM_RMAX + k) * HM_CL + i] : 0.0f;
    }
    __syncthreads();
    float acc[HM_RMAX][4];
#pragma unroll
    for (int k = 0; k < HM_RMAX; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[k][c] = 0.0f;
    for (int i0 = 0; i0 < szI; i0 += 8) {
        float v[8][4];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = min(i0 + u, szI - 1);
            const size_t off = row_off(i);
#pragma unroll
            for (int c = 0; c < 4; ++c) v[u][c] = ptr[c][off] * sgn[c];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (i0 + u >= szI) break;
#pragma unroll
            for (int k = 0; k < HM_RMAX; ++k)
                if (k < r) {
                    const float q = Qs[k][i0 + u];
#pragma unroll
                    for (int c = 0; c < 4; ++c) acc[k][c] = fmaf(q, v[u][c], acc[k][c]);
                }
        }
    }
    float mx = 0.0f;
#pragma unroll
    for (int k = 0; k < HM_RMAX; ++k)
        if (k < r)
#pragma unroll
            for (int c = 0; c < 4; ++c) mx = fmaxf(mx, fabsf(acc[k][c]));
    const float sc = pow2_at_least(lane63(wave_reduce63<true>(mx)));
    const float inv = 1.0f / sc;
    if (lane == 0) fscale[b] = sc;
#pragma unroll
    for (int k = 0; k < HM_RMAX; ++k)
        if (k < r) {
            hvec4 h;
#pragma unroll
            for (int c = 0; c < 4; ++c) h[c] = (_Float16)(acc[k][c] * inv);
            *(hvec4 *)(B16 + ((size_t)b * HM_RMAX + k) * HM_CL + 4 * lane) = h;
        }
}

// Phase 1 of the product: T[b][k] = sum_j B16[b][k][j] x_J(j) (unscaled) for every
// low-rank block b = (I, J); one wavefront per (column cluster J, chunk of row clusters),
// lane l holding x_J at the block columns 4 l .. 4 l + 3.
__global__ __launch_bounds__(64) void k_hm_gemv_t(const int *__restrict__ cl_idx,
                                                  const int *__restrict__ cl_size, int ncl,
                                                  int chunks, const uint8_t *__restrict__ rank,
                                                  const _Float16 *__restrict__ B16,
                                                  const float *__restrict__ x,
                                                  float *__restrict__ T)
{
    const int lane = threadIdx.x;
    const int J = blockIdx.x, ch = blockIdx.y;
    const int szJ = cl_size[J];
    fvec4 xv;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int p = 4 * lane + c;
        xv[c] = p < szJ ? x[cl_idx[(size_t)J * HM_CL + p]] : 0.0f;
    }
    // row clusters ch, ch + chunks, ...: the blocks of a chunk are spread over the row
    for (int I = ch; I < ncl; I += chunks) {
        const size_t b = (size_t)I * ncl + J;
        const int r = __builtin_amdgcn_readfirstlane(rank[b]);
        if (r == 0 || r == HM_DENSE) continue;
        const _Float16 *B = B16 + b * HM_RMAX * HM_CL + 4 * lane;
        float t[HM_RMAX];
#pragma unroll
        for (int k = 0; k < HM_RMAX; ++k)
            if (k < r) {
                const hvec4 h = *(const hvec4 *)(B + (size_t)k * HM_CL);
                t[k] = (float)h[0] * xv[0] + (float)h[1] * xv[1] + (float)h[2] * xv[2] +
                       (float)h[3] * xv[3];
            }
#pragma unroll
        for (int k = 0; k < HM_RMAX; ++k)
            if (k < r) t[k] = wave_reduce63<false>(t[k]);
        if (lane == 63)
#pragma unroll
            for (int k = 0; k < HM_RMAX; ++k)
                if (k < r) T[b * HM_RMAX + k] = t[k];
    }
}

// Phase 2: one wavefront per (row cluster I, chunk of column clusters), lane l holding the
// block rows 4 l .. 4 l + 3: ypart[chunk][row] = sum over the chunk's blocks of
// (A~_IJ x_J)(row), low-rank blocks as fscale Q16 T (8-byte loads of four rows of a Q
// column), dense ones row by row with wavefront sums.
__global__ __launch_bounds__(64) void k_hm_gemv(const int *__restrict__ cl_idx,
                                                const int *__restrict__ cl_size, int ncl,
                                                int chunks, const uint8_t *__restrict__ rank,
                                                const float *__restrict__ fscale,
                                                const _Float16 *__restrict__ Q16,
                                                const _Float16 *__restrict__ D16,
                                                const float *__restrict__ T,
                                                const float *__restrict__ x, size_t ldp,
                                                float *__restrict__ ypart)
{
    const int lane = threadIdx.x;
    const int I = blockIdx.x, ch = blockIdx.y;
    const int szI = cl_size[I];
    // column clusters ch, ch + chunks, ...: the dense blocks, which cluster around the
    // diagonal, are spread over the chunks
    fvec4 yacc = {0.f, 0.f, 0.f, 0.f};
    // low-rank blocks, two at a time (their loads in flight together)
    for (int J = ch; J < ncl; J += 2 * chunks) {
        fvec4 s2[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            s2[u] = fvec4{0.f, 0.f, 0.f, 0.f};
            const int Ju = J + u * chunks;
            if (Ju >= ncl) continue;
            const size_t b = (size_t)I * ncl + Ju;
            const int r = __builtin_amdgcn_readfirstlane(rank[b]);
            if (r == 0 || r == HM_DENSE) continue;
            const _Float16 *Q = Q16 + b * HM_RMAX * HM_CL + 4 * lane;
            const float *t = T + b * HM_RMAX;
            hvec4 q[HM_RMAX];
#pragma unroll
            for (int k = 0; k < HM_RMAX; ++k)
                if (k < r) q[k] = *(const hvec4 *)(Q + (size_t)k * HM_CL);
            fvec4 sv = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < HM_RMAX; ++k)
                if (k < r) {
                    const float tk = t[k];
#pragma unroll
                    for (int c = 0; c < 4; ++c) sv[c] = fmaf((float)q[k][c], tk, sv[c]);
                }
            s2[u] = sv * fscale[b];
        }
        yacc += s2[0] + s2[1];
    }
    // dense blocks, row by row: lane l holds x_J at the block columns 4 l .. 4 l + 3
    for (int J = ch; J < ncl; J += chunks) {
        const size_t b = (size_t)I * ncl + J;
        const int r = __builtin_amdgcn_readfirstlane(rank[b]);
        if (r != HM_DENSE) continue;
        const float fs = fscale[b];
        const int szJ = __builtin_amdgcn_readfirstlane(cl_size[J]);
        fvec4 xv;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int p = 4 * lane + c;
            xv[c] = p < szJ ? x[cl_idx[(size_t)J * HM_CL + p]] : 0.0f;
        }
        const _Float16 *D = D16 + b * HM_CL * HM_CL + 4 * lane;
        for (int i0 = 0; i0 < szI; i0 += 4) {
            float t[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const hvec4 h = *(const hvec4 *)(D + (size_t)min(i0 + u, szI - 1) * HM_CL);
                t[u] = (float)h[0] * xv[0] + (float)h[1] * xv[1] + (float)h[2] * xv[2] +
                       (float)h[3] * xv[3];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) t[u] = lane63(wave_reduce63<false>(t[u]));
            // rows i0 .. i0 + 3 belong to lane i0 / 4
            if (lane == (i0 >> 2))
#pragma unroll
                for (int u = 0; u < 4; ++u) yacc[u] = fmaf(t[u], fs, yacc[u]);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int li = 4 * lane + c;
        if (li < szI) ypart[(size_t)ch * ldp + cl_idx[(size_t)I * HM_CL + li]] = yacc[c];
    }
}

// y[i] = sum over chunks of ypart[c][i] (in chunk order) + diag[i] xd[i]
__global__ __launch_bounds__(256) void k_hm_finish(const float *__restrict__ ypart, size_t ldp,
                                                   int chunks, int n,
                                                   const double *__restrict__ diag,
                                                   const float *__restrict__ xd,
                                                   float *__restrict__ y)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < chunks; ++c) s += (double)ypart[(size_t)c * ldp + i];
    y[i] = (float)(s + diag[i] * (double)xd[i]);
}

int hm_chunks(int ncl)
{
    // about four waves per SIMD over the whole chip (256 CUs x 4 SIMDs) in the row phase
    const int c = (4096 + ncl - 1) / ncl;
    return std::max(1, std::min(64, c));
}

// row-cluster chunks of the T phase: about four waves per SIMD
static int hm_tchunks(int ncl) { return std::max(1, std::min(64, (4096 + ncl - 1) / ncl)); }

void launch_hm_build(const float *G, const float *H, size_t ld, int n, const uint8_t *gmask,
                     const int *cl_idx, const int *cl_size, const int *cl_seg, int ncl,
                     const uint8_t *near, float tol, float *Y, float *Mx, uint8_t *rank,
                     float *fscale, void *Q16, void *B16, void *D16, int agg, int nagg,
                     double *S, size_t ldS, hipStream_t s)
{
    hipLaunchKernelGGL(k_hm_sketch, dim3((n + 15) / 16), dim3(64), 0, s, G, H, ld, n, gmask,
                       (const int4 *)cl_seg, ncl, Y, Mx, agg, nagg, S, ldS);
    const unsigned nb = (unsigned)ncl * (unsigned)ncl;
    hipLaunchKernelGGL(k_hm_qr, dim3(nb), dim3(64), 0, s, Y, Mx, n, cl_idx, cl_size, ncl, near,
                       tol, rank, fscale, (_Float16 *)Q16);
    hipLaunchKernelGGL(k_hm_fill, dim3(nb), dim3(64), 0, s, G, H, ld, gmask, cl_idx, cl_size,
                       ncl, rank, fscale, (const _Float16 *)Q16, (_Float16 *)B16,
                       (_Float16 *)D16);
}

void launch_hm_gemv(const int *cl_idx, const int *cl_size, int ncl, const uint8_t *rank,
                    const float *fscale, const void *Q16, const void *B16, const void *D16,
                    int n, const float *x, const double *diag, const float *xd, float *T,
                    float *ypart, size_t ldp, float *y, hipStream_t s)
{
    const int ch = hm_chunks(ncl), tch = hm_tchunks(ncl);
    hipLaunchKernelGGL(k_hm_gemv_t, dim3(ncl, tch), dim3(64), 0, s, cl_idx, cl_size, ncl, tch,
                       rank, (const _Float16 *)B16, x, T);
    hipLaunchKernelGGL(k_hm_gemv, dim3(ncl, ch), dim3(64), 0, s, cl_idx, cl_size, ncl, ch,
                       rank, fscale, (const _Float16 *)Q16, (const _Float16 *)D16, T, x, ldp,
                       ypart);
    hipLaunchKernelGGL(k_hm_finish, dim3((n + 255) / 256), dim3(256), 0, s, ypart, ldp, ch, n,
                       diag, xd, y);
}

}  // namespace wb
