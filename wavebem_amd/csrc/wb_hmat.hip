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
// Product: k_hm_gather (x in cluster order), k_hm_gemv (one wavefront per row cluster and
// chunk of its work items: low-rank blocks and column parts of dense blocks) and
// k_hm_finish (chunk sums, the alpha diagonal).
#include <algorithm>
#include <functional>
#include <vector>

#include "wb_internal.h"

namespace wb {

namespace {

typedef _Float16 hvec4 __attribute__((ext_vector_type(4)));
typedef _Float16 hvec8 __attribute__((ext_vector_type(8)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef double dvec4 __attribute__((ext_vector_type(4)));

#ifndef WB_HM_BATCH
#define WB_HM_BATCH 16  // dense columns per batch of loads in the product
#endif

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

// Wavefront sums of 8 values at once (a transposing reduction): each step pairs a lane
// with a partner that holds the same set of value indices, keeps half of them and adds
// the partner's copy of that half (row_mirror splits on lane bit 3, row_half_mirror on
// bit 2, quad_perm [2,3,0,1] on bit 1); then the lane pairs, the rows and the row pairs
// are summed.  On return lane l holds the total of value ((l >> 1) & 7): 7 + 3 + 1 DPP
// adds and 2 cross-row shuffles where 8 separate sums take 8 x 6.
__device__ __forceinline__ float wave_sum8(const float (&p)[8])
{
    const unsigned l = threadIdx.x & 63u;
    const bool b3 = l & 8u, b2 = l & 4u, b1 = l & 2u;
    float a[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
        a[j] = (b3 ? p[j + 4] : p[j]) + WB_DPP_MOV(b3 ? p[j] : p[j + 4], 0x140);
    float c[2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
        c[j] = (b2 ? a[j + 2] : a[j]) + WB_DPP_MOV(b2 ? a[j] : a[j + 2], 0x141);
    float t = (b1 ? c[1] : c[0]) + WB_DPP_MOV(b1 ? c[0] : c[1], 0x4e);
    t += WB_DPP_MOV(t, 0xb1);
    t += __shfl_xor(t, 16);
    t += __shfl_xor(t, 32);
    return t;
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

// One wavefront per 16 consecutive device rows and a range of column clusters J, on the
// matrix cores (v_mfma_f32_16x16x4f32: A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k,
// D[i][j] in lane j + 16 (i / 4), register i % 4):
//   Y[J][row][k] = sum over the columns p of cluster J of A(row, col_J(p)) s_k(p),
//   Mx[J][row]   = max |A(row, col_J(p))|,
// lane (i, kk) reading row i at the positions q = 16 m + 4 kk + e of each segment (16 bytes
// per load).  The segment descriptors {first column, length, kind, unused} say whether a
// segment reads only G, only H or both (then column by column through gmask): pure
// segments go to separate G and H accumulators, combined as H - G at the end, so the
// values need no sign.  With S, also S[a][row] = the row's sum over the columns of
// aggregate a (agg consecutive device columns, agg >= HM_SEG, so a segment meets at most
// two aggregates), summed per segment in a fixed order.
__global__ __launch_bounds__(64) void k_hm_sketch(const float *__restrict__ G,
                                                  const float *__restrict__ H, size_t ld, int n,
                                                  const uint8_t *__restrict__ gmask,
                                                  const int4 *__restrict__ cl_seg, int ncl,
                                                  float *__restrict__ Y, float *__restrict__ Mx,
                                                  int agg, int nagg, double *__restrict__ S,
                                                  size_t ldS)
{
    // aggregate sums of the 16 rows, [a][row] (dynamic: 16 nagg floats with S)
    extern __shared__ float sacc[];
    const int lane = threadIdx.x, i = lane & 15, kk = lane >> 4;
    const int row0 = blockIdx.x * 16, row = row0 + i;
    const bool rok = row < n;
    const size_t off = (size_t)(rok ? row : 0) * ld;
    // clusters [J0, J1) of this wavefront: the row block's clusters are split over gridDim.y
    // wavefronts, whose aggregate sums go to parts S + y nagg ldS (added in order later)
    const int J0 = (int)((long)blockIdx.y * ncl / gridDim.y);
    const int J1 = (int)((long)(blockIdx.y + 1) * ncl / gridDim.y);
    if (S) S += (size_t)blockIdx.y * nagg * ldS;
    // the lane's sign of sketch vector i at position 16 g + 4 kk + e (g = 4 q + m)
    float sg[16][4];
#pragma unroll
    for (int g = 0; g < 16; ++g)
#pragma unroll
        for (int e = 0; e < 4; ++e)
            sg[g][e] = ((sign_bits((uint32_t)(16 * g + 4 * kk + e)) >> i) & 1u) ? 1.0f : -1.0f;
    if (S)
        for (int t = lane; t < 16 * nagg; t += 64) sacc[t] = 0.0f;
    __syncthreads();
    for (int J = J0; J < J1; ++J) {
        fvec4 aG0 = {0.f, 0.f, 0.f, 0.f}, aG1 = aG0, aH0 = aG0, aH1 = aG0;
        float mx = 0.0f;
#pragma unroll
        for (int q0 = 0; q0 < HM_SEGS; q0 += 2) {
            // two segments' loads in flight together
            int4 sd[2];
            fvec4 v[2][4];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                sd[h] = cl_seg[J * HM_SEGS + q0 + h];
                sd[h].x = __builtin_amdgcn_readfirstlane(sd[h].x);
                sd[h].y = __builtin_amdgcn_readfirstlane(sd[h].y);
                sd[h].z = __builtin_amdgcn_readfirstlane(sd[h].z);
                const int start = sd[h].x, len = sd[h].y, kind = sd[h].z;
                const float *src = (kind == 1 ? H : G) + off + start;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int p = 16 * m + 4 * kk;
                    fvec4 w = {0.f, 0.f, 0.f, 0.f};
                    if (rok && p < len) {
                        if (kind != 2) {
                            if (p + 4 <= len) {
                                __builtin_memcpy(&w, src + p, 16);
                            } else {
#pragma unroll
                                for (int e = 0; e < 4; ++e)
                                    if (p + e < len) w[e] = src[p + e];
                            }
                        } else {
                            // mixed: the operator's value (sign included), column by column
#pragma unroll
                            for (int e = 0; e < 4; ++e)
                                if (p + e < len) {
                                    const int col = start + p + e;
                                    const bool hh = (gmask[col >> 2] >> (col & 3)) & 1u;
                                    const float x = (hh ? H : G)[off + col];
                                    w[e] = hh ? x : -x;
                                }
                        }
                    }
                    v[h][m] = w;
                }
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int q = q0 + h;
                const int len = sd[h].y, kind = sd[h].z;
                if (len == 0) continue;
                float tot = 0.0f, lo = 0.0f;
                // columns below qb (segment position) belong to the segment's first aggregate
                const int a0 = S ? sd[h].x / agg : 0;
                const int qb = S ? (a0 + 1) * agg - sd[h].x : HM_SEG;
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const float x = v[h][m][e];
                        mx = fmaxf(mx, fabsf(x));
                        if (S) tot += x;
                        if (qb < HM_SEG) lo += (16 * m + 4 * kk + e < qb) ? x : 0.0f;
                    }
                // kind 0 reads G, whose operator entries are -G: to the G sums; kind 1 and
                // mixed segments (values with their sign) to the H sums
                if (kind == 0) {
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        aG0 = __builtin_amdgcn_mfma_f32_16x16x4f32(v[h][m][0], sg[4 * q + m][0], aG0, 0, 0, 0);
                        aG1 = __builtin_amdgcn_mfma_f32_16x16x4f32(v[h][m][1], sg[4 * q + m][1], aG1, 0, 0, 0);
                        aG0 = __builtin_amdgcn_mfma_f32_16x16x4f32(v[h][m][2], sg[4 * q + m][2], aG0, 0, 0, 0);
                        aG1 = __builtin_amdgcn_mfma_f32_16x16x4f32(v[h][m][3], sg[4 * q + m][3], aG1, 0, 0, 0);
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        aH0 = __builtin_amdgcn_mfma_f32_16x16x4f32(v[h][m][0], sg[4 * q + m][0], aH0, 0, 0, 0);
                        aH1 = __builtin_amdgcn_mfma_f32_16x16x4f32(v[h][m][1], sg[4 * q + m][1], aH1, 0, 0, 0);
                        aH0 = __builtin_amdgcn_mfma_f32_16x16x4f32(v[h][m][2], sg[4 * q + m][2], aH0, 0, 0, 0);
                        aH1 = __builtin_amdgcn_mfma_f32_16x16x4f32(v[h][m][3], sg[4 * q + m][3], aH1, 0, 0, 0);
                    }
                }
                if (S) {
                    if (qb >= HM_SEG) lo = tot;
                    const float sgn = kind == 0 ? -1.0f : 1.0f;
                    float hi = (tot - lo) * sgn;
                    lo *= sgn;
                    lo += __shfl_xor(lo, 16, 64);
                    lo += __shfl_xor(lo, 32, 64);
                    hi += __shfl_xor(hi, 16, 64);
                    hi += __shfl_xor(hi, 32, 64);
                    if (kk == 0) {
                        sacc[16 * a0 + i] += lo;
                        if (a0 + 1 < nagg) sacc[16 * (a0 + 1) + i] += hi;
                    }
                }
            }
        }
        const fvec4 acc = (aH0 + aH1) - (aG0 + aG1);
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
            if (r < n) S[(size_t)a * ldS + r] = (double)sacc[t];
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

// One wavefront per block: dense blocks -> D16 (column-major, 256 x 256, scaled by
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
                                                _Float16 *__restrict__ D16,
                                                const int *__restrict__ doff)
{
    __shared__ _Float16 Qs[HM_RMAX][HM_CL];  // Q in FP16, as stored (6 KB)
    const int b = blockIdx.x, lane = threadIdx.x;
    const int I = b / ncl, J = b % ncl;
    const int r = rank[b];
    if (r == 0) return;
    const int szI = cl_size[I], szJ = cl_size[J];
    // the lane's rows l, l + 64, ...: element offsets row * ld, read back with readlane
    uint64_t rowo[HM_CL / 64];
#pragma unroll
    for (int u = 0; u < HM_CL / 64; ++u) {
        const int li = lane + 64 * u;
        rowo[u] = (uint64_t)(li < szI ? cl_idx[(size_t)I * HM_CL + li] : 0) * ld;
    }
    const float *ptr[4];
    float sgn[4];
    bool same = true;  // four consecutive columns of one matrix: one 16-byte load per row
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int p = 4 * lane + c;
        if (p < szJ) {
            const int col = cl_idx[(size_t)J * HM_CL + p];
            const bool h = (gmask[col >> 2] >> (col & 3)) & 1u;
            ptr[c] = (h ? H : G) + col;
            sgn[c] = h ? 1.0f : -1.0f;
        } else {
            ptr[c] = G + c;  // a valid address of the row, weighted 0
            sgn[c] = 0.0f;
        }
        if (c > 0) same = same && ptr[c] == ptr[c - 1] + 1 && sgn[c] == sgn[0];
    }
    auto load4 = [&](size_t off, float (&v)[4]) {
        if (same) {
            fvec4 w;
            __builtin_memcpy(&w, ptr[0] + off, 16);
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = w[c] * sgn[0];
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = ptr[c][off] * sgn[c];
        }
    };
    auto row_off_u = [&](int u, int il) {
        const uint64_t v = rowo[u];
        const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, il);
        const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), il);
        return (size_t)(((uint64_t)hi << 32) | lo);
    };
    if (r == HM_DENSE) {
        // column-major: D16[b][col][row]; lane l writes its columns 4 l .. 4 l + 3 (zero
        // past |J|) for 8 rows at a time; rows past |I| are zero as well
        const float inv = 1.0f / fscale[b];
        _Float16 *Db = D16 + (size_t)doff[b] * HM_CL * HM_CL;
#pragma unroll
        for (int uq = 0; uq < HM_CL / 64; ++uq)
            for (int j0 = 0; j0 < 64; j0 += 8) {
                float v[8][4];
                if (64 * uq + j0 < szI) {
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        load4(row_off_u(uq, min(j0 + u, szI - 1 - 64 * uq)), v[u]);
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    hvec8 h;
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        h[u] = 64 * uq + j0 + u < szI ? (_Float16)(v[u][c] * inv)
                                                      : (_Float16)0.0f;
                    *(hvec8 *)(Db + (size_t)(4 * lane + c) * HM_CL + 64 * uq + j0) = h;
                }
            }
        return;
    }
    for (int e = lane; e < r * HM_CL; e += 64) {
        const int k = e / HM_CL, i = e % HM_CL;
        Qs[k][i] = i < szI ? Q16[((size_t)b * HM_RMAX + k) * HM_CL + i] : (_Float16)0.0f;
    }
    __syncthreads();
    float acc[HM_RMAX][4];
#pragma unroll
    for (int k = 0; k < HM_RMAX; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[k][c] = 0.0f;
#pragma unroll
    for (int uq = 0; uq < HM_CL / 64; ++uq)
        for (int j0 = 0; j0 < 64 && 64 * uq + j0 < szI; j0 += 8) {
            float v[8][4];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                load4(row_off_u(uq, min(j0 + u, szI - 1 - 64 * uq)), v[u]);
            const int i0 = 64 * uq + j0;
#pragma unroll
            for (int k = 0; k < HM_RMAX; ++k)
                if (k < r) {
                    // Q rows i0 .. i0 + 7 (zero past the block's rows)
                    const hvec8 qh = *(const hvec8 *)&Qs[k][i0];
                    const fvec4 q0 = {(float)qh[0], (float)qh[1], (float)qh[2], (float)qh[3]};
                    const fvec4 q1 = {(float)qh[4], (float)qh[5], (float)qh[6], (float)qh[7]};
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        float t = acc[k][c];
                        t = fmaf(q0[0], v[0][c], t);
                        t = fmaf(q0[1], v[1][c], t);
                        t = fmaf(q0[2], v[2][c], t);
                        t = fmaf(q0[3], v[3][c], t);
                        t = fmaf(q1[0], v[4][c], t);
                        t = fmaf(q1[1], v[5][c], t);
                        t = fmaf(q1[2], v[6][c], t);
                        t = fmaf(q1[3], v[7][c], t);
                        acc[k][c] = t;
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

// Phase 1 of the product: x in cluster order, xc[J][p] = x(cl_idx[J][p]) (0 past |J|), so
// that the work items read x_J with one contiguous load.
__global__ __launch_bounds__(256) void k_hm_gather(const int *__restrict__ cl_idx,
                                                   const int *__restrict__ cl_size,
                                                   const float *__restrict__ x,
                                                   float *__restrict__ xc)
{
    const int J = blockIdx.x, p = threadIdx.x;
    xc[(size_t)J * HM_CL + p] = p < cl_size[J] ? x[cl_idx[(size_t)J * HM_CL + p]] : 0.0f;
}

// Phase 2: one wavefront per (row cluster I, chunk ch), lane l holding the block rows
// 4 l .. 4 l + 3: ypart[ch][row] = sum over the chunk's work items of (A~_IJ x_J)(row).
// The items (built on the host from the ranks, hm_plan_items) are the low-rank blocks of
// row cluster I and HM_DCOLS-column parts of its dense blocks.  A low-rank block is
// applied as fscale Q16 (B16 x_J): lane l forms its part of each of the r sums
// t_k = B16[k] . x_J from the block columns 4 l .. 4 l + 3, the wavefront sums them (DPP)
// and every lane adds q_k(rows) t_k; the B16 and Q16 loads (8 bytes per lane and column)
// are issued together.  Dense parts are streamed column by column in batches of
// WB_HM_BATCH columns (independent 512-byte loads per wavefront), so no wavefront walks a
// whole dense block alone.
__global__ __launch_bounds__(64) void k_hm_gemv(const int *__restrict__ cl_idx,
                                                const int *__restrict__ cl_size, int ncl,
                                                int chunks, const uint8_t *__restrict__ rank,
                                                const float *__restrict__ fscale,
                                                const _Float16 *__restrict__ Q16,
                                                const _Float16 *__restrict__ B16,
                                                const _Float16 *__restrict__ D16,
                                                const int *__restrict__ doff,
                                                const float *__restrict__ xc, size_t ldp,
                                                float *__restrict__ ypart,
                                                const int *__restrict__ item_ptr,
                                                const int *__restrict__ items)
{
    const int lane = threadIdx.x;
    const int I = blockIdx.x, ch = blockIdx.y;
    const int szI = cl_size[I];
    const int p0 = __builtin_amdgcn_readfirstlane(item_ptr[I * chunks + ch]);
    const int p1 = __builtin_amdgcn_readfirstlane(item_ptr[I * chunks + ch + 1]);
    fvec4 yacc = {0.f, 0.f, 0.f, 0.f};
    auto lowrank = [&](int J) {
        const size_t b = (size_t)I * ncl + J;
        const int r = __builtin_amdgcn_readfirstlane(rank[b]);
        const fvec4 xv = *(const fvec4 *)(xc + (size_t)J * HM_CL + 4 * lane);
        const _Float16 *Bp = B16 + b * HM_RMAX * HM_CL + 4 * lane;
        const _Float16 *Q = Q16 + b * HM_RMAX * HM_CL + 4 * lane;
        hvec4 bk[HM_RMAX], q[HM_RMAX];
#pragma unroll
        for (int k = 0; k < HM_RMAX; ++k)
            if (k < r) {
                bk[k] = *(const hvec4 *)(Bp + (size_t)k * HM_CL);
                q[k] = *(const hvec4 *)(Q + (size_t)k * HM_CL);
            }
        fvec4 sv = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < HM_RMAX; ++k)
            if (k < r) {
                const float part = (float)bk[k][0] * xv[0] + (float)bk[k][1] * xv[1] +
                                   (float)bk[k][2] * xv[2] + (float)bk[k][3] * xv[3];
                const float tk = lane63(wave_reduce63<false>(part));
#pragma unroll
                for (int c = 0; c < 4; ++c) sv[c] = fmaf((float)q[k][c], tk, sv[c]);
            }
        return sv * fscale[b];
    };
    int p = p0;
    // low-rank items (listed first)
    for (; p < p1; ++p) {
        const int it = __builtin_amdgcn_readfirstlane(items[p]);
        if ((it & HM_ITEM_MASK) != HM_ITEM_LOWRANK) break;
        yacc += lowrank(it >> HM_ITEM_SHIFT);
    }
    // dense parts: columns HM_DCOLS q .. HM_DCOLS (q + 1) - 1 of block (I, J), column-major
    // storage; lanes 0 .. HM_DCOLS / 4 - 1 hold x_J at the part's columns 4 l .. 4 l + 3
    // for the broadcast
    constexpr int NB = WB_HM_BATCH;
    static_assert(HM_DCOLS % NB == 0 && NB % 4 == 0 && HM_DCOLS <= 4 * 64, "batch");
    static_assert(HM_CL / HM_DCOLS < HM_ITEM_LOWRANK, "item kinds");
    for (; p < p1; ++p) {
        const int it = __builtin_amdgcn_readfirstlane(items[p]);
        const int J = it >> HM_ITEM_SHIFT, q = it & HM_ITEM_MASK;
        const size_t b = (size_t)I * ncl + J;
        const float fs = fscale[b];
        const int szJ = __builtin_amdgcn_readfirstlane(cl_size[J]);
        const int cb = HM_DCOLS * q, ce = min(cb + HM_DCOLS, szJ);
        // x of the part's columns: wave-uniform addresses, so scalar loads
        const float *xp = xc + (size_t)J * HM_CL;
        const _Float16 *D = D16 + (size_t)doff[b] * HM_CL * HM_CL + 4 * lane;
        fvec4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = a0;
        for (int c0 = cb; c0 < ce; c0 += NB) {
            // columns c0 .. c0 + NB - 1 (inside the part and the 256 stored columns, zero past
            // |J|; x is zero there too)
            hvec4 h[NB];
#pragma unroll
            for (int u = 0; u < NB; ++u) h[u] = *(const hvec4 *)(D + (size_t)(c0 + u) * HM_CL);
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const float xcol = xp[c0 + u];
                fvec4 &a = (u & 1) ? a1 : a0;
#pragma unroll
                for (int e = 0; e < 4; ++e) a[e] = fmaf((float)h[u][e], xcol, a[e]);
            }
        }
        yacc += (a0 + a1) * fs;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int li = 4 * lane + c;
        if (li < szI) ypart[(size_t)ch * ldp + cl_idx[(size_t)I * HM_CL + li]] = yacc[c];
    }
}

// ------------------------------------------------------------------ several vectors
// The product for up to HM_KV vectors at once, over the same work items: the operator is
// read once per launch.  xcm holds the vectors in cluster order ([v][J HM_CL + p]); a
// low-rank item forms t_{k,v} = B_k . x_J^v for every vector (lane partials, one
// transposing wavefront sum per k for the 8 vectors) and adds Q t; a dense part multiplies
// every streamed column into every vector, x of its columns read with scalar loads.
// ypart is [ch][v][row].
constexpr int HM_KV = 8;
static_assert(HM_KV == 8, "the low-rank sums use wave_sum8");

__global__ __launch_bounds__(HM_CL) void k_hm_gather_multi(const int *__restrict__ cl_idx,
                                                           const int *__restrict__ cl_size,
                                                           int ncl, const float *__restrict__ X,
                                                           size_t ldx, float *__restrict__ xcm)
{
    const int J = blockIdx.x, v = blockIdx.y, p = threadIdx.x;
    xcm[((size_t)v * ncl + J) * HM_CL + p] =
        p < cl_size[J] ? X[(size_t)v * ldx + cl_idx[(size_t)J * HM_CL + p]] : 0.0f;
}

__global__ __launch_bounds__(64) void k_hm_gemv_multi(
    const int *__restrict__ cl_idx, const int *__restrict__ cl_size, int ncl, int chunks,
    const uint8_t *__restrict__ rank, const float *__restrict__ fscale,
    const _Float16 *__restrict__ Q16, const _Float16 *__restrict__ B16,
    const _Float16 *__restrict__ D16, const int *__restrict__ doff,
    const float *__restrict__ xcm, int k, size_t ldp, float *__restrict__ ypart,
    const int *__restrict__ item_ptr, const int *__restrict__ items)
{
    const int lane = threadIdx.x;
    const int I = blockIdx.x, ch = blockIdx.y;
    const int szI = cl_size[I];
    const size_t vs = (size_t)ncl * HM_CL;  // vector stride of xcm
    const int p0 = __builtin_amdgcn_readfirstlane(item_ptr[I * chunks + ch]);
    const int p1 = __builtin_amdgcn_readfirstlane(item_ptr[I * chunks + ch + 1]);
    fvec4 yacc[HM_KV];
#pragma unroll
    for (int v = 0; v < HM_KV; ++v) yacc[v] = fvec4{0.f, 0.f, 0.f, 0.f};
    constexpr int NB = WB_HM_BATCH;
    for (int p = p0; p < p1; ++p) {
        const int it = __builtin_amdgcn_readfirstlane(items[p]);
        const int J = it >> HM_ITEM_SHIFT, kind = it & HM_ITEM_MASK;
        const size_t b = (size_t)I * ncl + J;
        const float fs = fscale[b];
        if (kind == HM_ITEM_LOWRANK) {
            const int r = __builtin_amdgcn_readfirstlane(rank[b]);
            const _Float16 *Bp = B16 + b * HM_RMAX * HM_CL + 4 * lane;
            const _Float16 *Q = Q16 + b * HM_RMAX * HM_CL + 4 * lane;
            hvec4 bk[HM_RMAX], q[HM_RMAX];
#pragma unroll
            for (int kr = 0; kr < HM_RMAX; ++kr)
                if (kr < r) {
                    bk[kr] = *(const hvec4 *)(Bp + (size_t)kr * HM_CL);
                    q[kr] = *(const hvec4 *)(Q + (size_t)kr * HM_CL);
                }
            fvec4 xv[HM_KV];
#pragma unroll
            for (int v = 0; v < HM_KV; ++v)
                xv[v] = v < k ? *(const fvec4 *)(xcm + v * vs + (size_t)J * HM_CL + 4 * lane)
                              : fvec4{0.f, 0.f, 0.f, 0.f};
            // t_{kr,v} for the 8 vectors in one transposing sum (x is zero past k)
#pragma unroll
            for (int kr = 0; kr < HM_RMAX; ++kr)
                if (kr < r) {
                    float part[HM_KV];
#pragma unroll
                    for (int v = 0; v < HM_KV; ++v)
                        part[v] = (float)bk[kr][0] * xv[v][0] + (float)bk[kr][1] * xv[v][1] +
                                  (float)bk[kr][2] * xv[v][2] + (float)bk[kr][3] * xv[v][3];
                    const float ts = wave_sum8(part);
#pragma unroll
                    for (int v = 0; v < HM_KV; ++v)
                        if (v < k) {
                            const float t =
                                __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ts), 2 * v)) *
                                fs;
#pragma unroll
                            for (int c = 0; c < 4; ++c)
                                yacc[v][c] = fmaf((float)q[kr][c], t, yacc[v][c]);
                        }
                }
        } else {
            const int szJ = __builtin_amdgcn_readfirstlane(cl_size[J]);
            const int cb = HM_DCOLS * kind, ce = min(cb + HM_DCOLS, szJ);
            // x of the part's columns: wave-uniform addresses, so scalar loads
            const float *xp = xcm + (size_t)J * HM_CL;
            const _Float16 *D = D16 + (size_t)doff[b] * HM_CL * HM_CL + 4 * lane;
            fvec4 a[HM_KV];
#pragma unroll
            for (int v = 0; v < HM_KV; ++v) a[v] = fvec4{0.f, 0.f, 0.f, 0.f};
            for (int c0 = cb; c0 < ce; c0 += NB) {
                hvec4 h[NB];
#pragma unroll
                for (int u = 0; u < NB; ++u) h[u] = *(const hvec4 *)(D + (size_t)(c0 + u) * HM_CL);
#pragma unroll
                for (int v = 0; v < HM_KV; ++v)
                    if (v < k)
#pragma unroll
                        for (int u = 0; u < NB; ++u) {
                            const float xcol = xp[v * vs + c0 + u];
#pragma unroll
                            for (int e = 0; e < 4; ++e)
                                a[v][e] = fmaf((float)h[u][e], xcol, a[v][e]);
                        }
            }
#pragma unroll
            for (int v = 0; v < HM_KV; ++v)
                if (v < k) yacc[v] += a[v] * fs;
        }
    }
#pragma unroll
    for (int v = 0; v < HM_KV; ++v)
        if (v < k)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int li = 4 * lane + c;
                if (li < szI)
                    ypart[((size_t)ch * k + v) * ldp + cl_idx[(size_t)I * HM_CL + li]] =
                        yacc[v][c];
            }
}

// Y[v][i] = sum over chunks of ypart[c][v][i] (as k_hm_finish) + diag[i] X[v][i]; with
// line, constrained rows get X[v][r] - sum w X[v][col] instead
__global__ __launch_bounds__(256) void k_hm_finish_multi(
    const float *__restrict__ ypart, size_t ldp, int chunks, int n, int k,
    const double *__restrict__ diag, const float *__restrict__ X, size_t ldx,
    float *__restrict__ Y, size_t ldy, const int *__restrict__ line,
    const int *__restrict__ cptr, const int *__restrict__ ccols, const double *__restrict__ cw)
{
    __shared__ double qs[4][64];
    const int rl = threadIdx.x & 63, q = threadIdx.x >> 6, v = blockIdx.y;
    const int i = blockIdx.x * 64 + rl;
    const int c0 = (int)((long)q * chunks / 4), c1 = (int)((long)(q + 1) * chunks / 4);
    double s = 0.0;
    if (i < n)
        for (int c = c0; c < c1; ++c) s += (double)ypart[((size_t)c * k + v) * ldp + i];
    qs[q][rl] = s;
    __syncthreads();
    if (q != 0 || i >= n) return;
    const float *x = X + (size_t)v * ldx;
    float yv = (float)((((qs[0][rl] + qs[1][rl]) + qs[2][rl]) + qs[3][rl]) +
                       diag[i] * (double)x[i]);
    if (line) {
        const int l = line[i];
        if (l >= 0) {
            double t = (double)x[i];
            for (int p = cptr[l]; p < cptr[l + 1]; ++p) t -= cw[p] * (double)x[ccols[p]];
            yv = (float)t;
        }
    }
    Y[(size_t)v * ldy + i] = yv;
}

// y[i] = sum over chunks of ypart[c][i] + diag[i] xd[i]: 64 rows per block, the chunks
// split into four contiguous quarters summed by four threads of a row (in chunk order),
// the quarters then added in order, so the sum is the same on every call.
// With line (the constraint line of each row, -1 for free rows): constrained rows get the
// line's value instead, x_i - sum_p w_p x_{col_p} (as k_constraint_rows).  With Cpart:
// the block's sums of y over the (at most two, agg >= 64) aggregates its rows touch,
// Cpart[2 b + k] for aggregate (64 b) / agg + k, added in a fixed (butterfly) order.
__global__ __launch_bounds__(256) void k_hm_finish(const float *__restrict__ ypart, size_t ldp,
                                                   int chunks, int n,
                                                   const double *__restrict__ diag,
                                                   const float *__restrict__ xd,
                                                   float *__restrict__ y,
                                                   const int *__restrict__ line,
                                                   const int *__restrict__ cptr,
                                                   const int *__restrict__ ccols,
                                                   const double *__restrict__ cw,
                                                   const float *__restrict__ x, int agg,
                                                   double *__restrict__ Cpart)
{
    __shared__ double qs[4][64];
    const int rl = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + rl;
    const int c0 = (int)((long)q * chunks / 4), c1 = (int)((long)(q + 1) * chunks / 4);
    double s = 0.0;
    if (i < n)
        for (int c = c0; c < c1; ++c) s += (double)ypart[(size_t)c * ldp + i];
    qs[q][rl] = s;
    __syncthreads();
    if (q != 0) return;
    float yv = 0.0f;
    if (i < n) {
        yv = (float)((((qs[0][rl] + qs[1][rl]) + qs[2][rl]) + qs[3][rl]) +
                     diag[i] * (double)xd[i]);
        if (line) {
            const int l = line[i];
            if (l >= 0) {
                double v = (double)x[i];
                for (int p = cptr[l]; p < cptr[l + 1]; ++p) v -= cw[p] * (double)x[ccols[p]];
                yv = (float)v;
            }
        }
        y[i] = yv;
    }
    if (Cpart) {
        const int a0 = (blockIdx.x * 64) / agg, a = i < n ? i / agg : -1;
        double s0 = a == a0 ? (double)yv : 0.0, s1 = a == a0 + 1 ? (double)yv : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            s0 += __shfl_xor(s0, off, 64);
            s1 += __shfl_xor(s1, off, 64);
        }
        if (rl == 0) {
            Cpart[2 * blockIdx.x] = s0;
            Cpart[2 * blockIdx.x + 1] = s1;
        }
    }
}

// S[a][r] += the other sketch parts' sums (parts 1 .. nparts - 1, in order)
__global__ __launch_bounds__(256) void k_hm_sum_parts(int n, int nagg, int nparts, size_t ldS,
                                                      double *__restrict__ S)
{
    const int r = blockIdx.x * 256 + threadIdx.x, a = blockIdx.y;
    if (r >= n) return;
    double v = S[(size_t)a * ldS + r];
    for (int p = 1; p < nparts; ++p) v += S[((size_t)p * nagg + a) * ldS + r];
    S[(size_t)a * ldS + r] = v;
}

// wavefronts of each product phase (1024 SIMDs: WB_HM_WAVES / 1024 per SIMD)
#ifndef WB_HM_WAVES
#define WB_HM_WAVES 8192
#endif
int hm_chunks(int ncl)
{
    const int c = (WB_HM_WAVES + ncl - 1) / ncl;
    return std::max(1, std::min(128, c));
}


void launch_hm_rank(const float *G, const float *H, size_t ld, int n, const uint8_t *gmask,
                    const int *cl_idx, const int *cl_size, const int *cl_seg, int ncl,
                    const uint8_t *near, float tol, float *Y, float *Mx, uint8_t *rank,
                    float *fscale, void *Q16, int agg, int nagg, double *S, size_t ldS,
                    hipStream_t s)
{
    const size_t lds = S ? sizeof(float) * 16 * (size_t)std::max(nagg, 1) : 0;
    hipLaunchKernelGGL(k_hm_sketch, dim3((n + 15) / 16, HM_SKETCH_PARTS), dim3(64), lds, s, G, H,
                       ld, n, gmask, (const int4 *)cl_seg, ncl, Y, Mx, agg, nagg, S, ldS);
    if (S && HM_SKETCH_PARTS > 1)
        hipLaunchKernelGGL(k_hm_sum_parts, dim3((n + 255) / 256, nagg), dim3(256), 0, s, n, nagg,
                           HM_SKETCH_PARTS, ldS, S);
    const unsigned nb = (unsigned)ncl * (unsigned)ncl;
    hipLaunchKernelGGL(k_hm_qr, dim3(nb), dim3(64), 0, s, Y, Mx, n, cl_idx, cl_size, ncl, near,
                       tol, rank, fscale, (_Float16 *)Q16);
}

void launch_hm_fill(const float *G, const float *H, size_t ld, const uint8_t *gmask,
                    const int *cl_idx, const int *cl_size, int ncl, const uint8_t *rank,
                    float *fscale, const void *Q16, void *B16, void *D16, const int *doff,
                    hipStream_t s)
{
    const unsigned nb = (unsigned)ncl * (unsigned)ncl;
    hipLaunchKernelGGL(k_hm_fill, dim3(nb), dim3(64), 0, s, G, H, ld, gmask, cl_idx, cl_size,
                       ncl, rank, fscale, (const _Float16 *)Q16, (_Float16 *)B16,
                       (_Float16 *)D16, doff);
}

void hm_plan_items(int ncl, const uint8_t *rank, const int *cl_size, std::vector<int> &item_ptr,
                   std::vector<int> &items)
{
    const int chunks = hm_chunks(ncl);
    item_ptr.assign((size_t)ncl * chunks + 1, 0);
    items.clear();
    std::vector<int> row;                 // the row cluster's items: low-rank, then dense
    std::vector<int> weight, chunk_of, order, count(chunks + 1);
    std::vector<uint64_t> heap(chunks);  // (load << 8) | chunk, a min-heap
    const int w_lowrank = 8 + 16 * HM_RMAX, w_dense = 8 * HM_DCOLS;  // weights in 64-byte units
    const int wmax = std::max(w_lowrank, w_dense);
    std::vector<int> cnt(wmax + 2);
    for (int I = 0; I < ncl; ++I) {
        row.clear();
        weight.clear();
        for (int J = 0; J < ncl; ++J) {
            const int r = rank[(size_t)I * ncl + J];
            if (r == 0 || r == HM_DENSE) continue;
            row.push_back((J << HM_ITEM_SHIFT) | HM_ITEM_LOWRANK);
            weight.push_back(8 + 16 * r);  // r columns of B and of Q, 512 bytes each
        }
        for (int J = 0; J < ncl; ++J)
            if (rank[(size_t)I * ncl + J] == HM_DENSE)
                for (int q = 0; HM_DCOLS * q < cl_size[J]; ++q) {
                    row.push_back((J << HM_ITEM_SHIFT) | q);
                    weight.push_back(w_dense);  // HM_DCOLS columns of 512 bytes
                }
        const int ni = (int)row.size();
        // largest weight first (a stable counting sort over the weights 1 .. wmax)
        order.resize(ni);
        {
            std::fill(cnt.begin(), cnt.end(), 0);
            for (int k = 0; k < ni; ++k) ++cnt[wmax - weight[k] + 1];
            for (int w = 1; w <= wmax + 1; ++w) cnt[w] += cnt[w - 1];
            for (int k = 0; k < ni; ++k) order[cnt[wmax - weight[k]]++] = k;
        }
        // each to the least loaded chunk, lowest index on ties
        for (int k = 0; k < chunks; ++k) heap[k] = (uint64_t)k;
        std::make_heap(heap.begin(), heap.end(), std::greater<uint64_t>());
        chunk_of.resize(ni);
        for (int k : order) {
            std::pop_heap(heap.begin(), heap.end(), std::greater<uint64_t>());
            uint64_t &top = heap.back();
            const int c = (int)(top & 0xff);
            chunk_of[k] = c;
            top = ((top >> 8) + (uint64_t)weight[k]) << 8 | (uint64_t)c;
            std::push_heap(heap.begin(), heap.end(), std::greater<uint64_t>());
        }
        // grouped by chunk, keeping the row's order (low-rank first, each kind by column)
        std::fill(count.begin(), count.end(), 0);
        for (int k = 0; k < ni; ++k) ++count[chunk_of[k] + 1];
        for (int c = 0; c < chunks; ++c) count[c + 1] += count[c];
        const size_t base = items.size();
        items.resize(base + ni);
        for (int c = 0; c < chunks; ++c) item_ptr[(size_t)I * chunks + c + 1] = (int)(base + count[c + 1]);
        for (int k = 0; k < ni; ++k) items[base + count[chunk_of[k]]++] = row[k];
    }
    if (items.empty()) items.push_back(0);
}

void launch_hm_gemv(const int *cl_idx, const int *cl_size, int ncl, const uint8_t *rank,
                    const float *fscale, const void *Q16, const void *B16, const void *D16,
                    const int *doff, const int *item_ptr, const int *items, int n,
                    const float *x, const double *diag, const float *xd, float *xc, float *ypart,
                    size_t ldp, float *y, hipStream_t s, bool xc_ready, const int *line,
                    const int *cptr, const int *ccols, const double *cw, int agg,
                    double *Cpart)
{
    const int ch = hm_chunks(ncl);
    if (!xc_ready)
        hipLaunchKernelGGL(k_hm_gather, dim3(ncl), dim3(HM_CL), 0, s, cl_idx, cl_size, x, xc);
    hipLaunchKernelGGL(k_hm_gemv, dim3(ncl, ch), dim3(64), 0, s, cl_idx, cl_size, ncl, ch,
                       rank, fscale, (const _Float16 *)Q16, (const _Float16 *)B16,
                       (const _Float16 *)D16, doff, xc, ldp, ypart, item_ptr, items);
    hipLaunchKernelGGL(k_hm_finish, dim3((n + 63) / 64), dim3(256), 0, s, ypart, ldp, ch, n,
                       diag, xd, y, line, cptr, ccols, cw, x, agg, Cpart);
}

size_t hm_multi_work_floats(int ncl, size_t ldp, size_t *xc_floats)
{
    if (xc_floats) *xc_floats = (size_t)HM_KV * ncl * HM_CL;
    return (size_t)hm_chunks(ncl) * HM_KV * ldp;
}

void launch_hm_gemv_multi(const int *cl_idx, const int *cl_size, int ncl, const uint8_t *rank,
                          const float *fscale, const void *Q16, const void *B16,
                          const void *D16, const int *doff, const int *item_ptr,
                          const int *items, int n, int k, const float *X, size_t ldx,
                          const double *diag, float *xcm, float *ypart, size_t ldp, float *Y,
                          size_t ldy, const int *line, const int *cptr, const int *ccols,
                          const double *cw, hipStream_t s)
{
    const int ch = hm_chunks(ncl);
    for (int g = 0; g < k; g += HM_KV) {
        const int kg = std::min(HM_KV, k - g);
        const float *Xg = X + (size_t)g * ldx;
        hipLaunchKernelGGL(k_hm_gather_multi, dim3(ncl, kg), dim3(HM_CL), 0, s, cl_idx, cl_size,
                           ncl, Xg, ldx, xcm);
        hipLaunchKernelGGL(k_hm_gemv_multi, dim3(ncl, ch), dim3(64), 0, s, cl_idx, cl_size, ncl,
                           ch, rank, fscale, (const _Float16 *)Q16, (const _Float16 *)B16,
                           (const _Float16 *)D16, doff, xcm, kg, ldp, ypart, item_ptr, items);
        hipLaunchKernelGGL(k_hm_finish_multi, dim3((n + 63) / 64, kg), dim3(256), 0, s, ypart,
                           ldp, ch, n, kg, diag, Xg, ldx, Y + (size_t)g * ldy, ldy, line, cptr,
                           ccols, cw);
    }
}

}  // namespace wb
