// wb_multi.hip -- several right-hand sides at once (MFMA) on gfx950.
//
// Y[:, v] = A_sel X[:, v] (+ diag .* Xd[:, v]) for v < k <= 16 (X, Xd with column stride
// ldx, Y with ldy): the column-selected operator of BEMProblem::vmult / compute_rhs
// (source/bem_problem.cc:621-717) applied to k vectors in one pass over the matrix.  The
// product is a dense contraction, so it runs on the matrix cores: v_mfma_f32_16x16x4f32
// for the FP32 (hi) matrix, v_mfma_f64_16x16x4f64 for the FP64 matrix and for hi + lo
// (converted to FP64 on load).  Operand layout (checked on the device): A[i][kk] in lane
// i + 16 kk and B[kk][j] in lane j + 16 kk for both; the result D[i][j] is in lane
// j + 16 (i / 4), register i % 4 for FP32 but in lane j + 16 (i % 4), register i / 4 for
// FP64.
//
// Each wavefront owns two 16-row tiles: the right-hand-side operand of a step is loaded
// once for both, and the two accumulator chains are independent, so the matrix cores do
// not wait on one another's results.  The four lanes of a row (k-slots) read adjacent
// 4-column groups, so a row segment of 64 (FP32) or 128 (FP64) bytes comes from one load
// instruction.
#include <algorithm>

#include "wb_internal.h"

namespace wb {

namespace {

typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef unsigned short uvec4 __attribute__((ext_vector_type(4)));

constexpr int MW = 4;   // wavefronts per block
constexpr int RT = 2;   // 16-row tiles per wavefront
constexpr int RPB = 16 * RT * MW;  // rows per block

__device__ __forceinline__ int tile_row(int slot, const int *rows, int n_rows)
{
    return slot < n_rows ? (rows ? rows[slot] : slot) : (rows ? rows[0] : 0);
}

// FP32 hi matrix, float vectors; FP32 matrix-core sums moved into FP64 every 128 columns,
// so rounding grows with the 128-column blocks, not with n.
__global__ __launch_bounds__(64 * MW) void k_gemm_sel_f32(
    const float *__restrict__ G, const float *__restrict__ H, size_t ld,
    const int4 *__restrict__ segs, int n_segs, const uint8_t *__restrict__ gmask, int flip,
    float s_sel, float s_uns, int k, const float *__restrict__ X, size_t ldx,
    const int *__restrict__ rows, int n_rows, int groups_per_chunk, double *__restrict__ part)
{
    const int gc0 = blockIdx.y * groups_per_chunk, gc1 = gc0 + groups_per_chunk;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int base = (blockIdx.x * MW + wave) * 16 * RT;
    const int i = lane & 15, q = lane >> 4;  // row in tile, k-slot
    size_t roff[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) roff[t] = (size_t)tile_row(base + 16 * t + i, rows, n_rows) * ld;
    const int v = lane & 15;  // right-hand side fed by this lane's B operand
    const bool vok = v < k;
    const float *Xv = X + (size_t)(vok ? v : 0) * ldx;
    fvec4 c[RT];
    double cd[RT][4];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        c[t] = fvec4{0.f, 0.f, 0.f, 0.f};
        cd[t][0] = cd[t][1] = cd[t][2] = cd[t][3] = 0.0;
    }
    int since = 0;
    for (int si = 0; si < n_segs; ++si) {
        const int4 sg = segs[si];
        const int g_lo = max(sg.x, gc0), g_hi = min(sg.y, gc1);  // this chunk's part
        if (g_lo >= g_hi) continue;
        const int c_end = 4 * g_hi;
        const unsigned fm = flip ? 0xfu : 0u;
        const bool pure = sg.z < 2;
        const bool use_h = (sg.z ^ flip) != 0;
        for (int cb = 4 * g_lo; cb < c_end; cb += 32) {
            if (++since == 4) {
#pragma unroll
                for (int t = 0; t < RT; ++t) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) cd[t][r] += (double)c[t][r];
                    c[t] = fvec4{0.f, 0.f, 0.f, 0.f};
                }
                since = 0;
            }
            float a[RT][8], b[8];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                // the four k-slots of a row read 64 contiguous bytes per load instruction
                const int cc = cb + 16 * h + 4 * q;
                fvec4 av[RT], xv = {0.f, 0.f, 0.f, 0.f};
                float sgn[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int t = 0; t < RT; ++t) av[t] = fvec4{0.f, 0.f, 0.f, 0.f};
                if (cc < c_end) {
                    const int g = cc >> 2;
                    if (pure) {
                        const float *M = use_h ? H : G;
#pragma unroll
                        for (int t = 0; t < RT; ++t)
                            av[t] = __builtin_nontemporal_load((const fvec4 *)(M + roff[t]) + g);
                        const float s = use_h ? s_sel : s_uns;
                        sgn[0] = sgn[1] = sgn[2] = sgn[3] = s;
                    } else {
                        const unsigned m = gmask[g] ^ fm;
#pragma unroll
                        for (int t = 0; t < RT; ++t) {
                            const fvec4 ga = *((const fvec4 *)(G + roff[t]) + g);
                            const fvec4 ha = *((const fvec4 *)(H + roff[t]) + g);
                            av[t].x = (m & 1) ? ha.x : ga.x;
                            av[t].y = (m & 2) ? ha.y : ga.y;
                            av[t].z = (m & 4) ? ha.z : ga.z;
                            av[t].w = (m & 8) ? ha.w : ga.w;
                        }
#pragma unroll
                        for (int e = 0; e < 4; ++e) sgn[e] = ((m >> e) & 1) ? s_sel : s_uns;
                    }
                    if (vok) xv = *((const fvec4 *)Xv + g);
                }
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    a[t][4 * h + 0] = av[t].x;
                    a[t][4 * h + 1] = av[t].y;
                    a[t][4 * h + 2] = av[t].z;
                    a[t][4 * h + 3] = av[t].w;
                }
                b[4 * h + 0] = xv.x * sgn[0];
                b[4 * h + 1] = xv.y * sgn[1];
                b[4 * h + 2] = xv.z * sgn[2];
                b[4 * h + 3] = xv.w * sgn[3];
            }
            // step s = 4h + e contracts column cb + 16h + 4q + e of every lane's k-slot q
#pragma unroll
            for (int s = 0; s < 8; ++s)
#pragma unroll
                for (int t = 0; t < RT; ++t)
                    c[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][s], b[s], c[t], 0, 0, 0);
        }
    }
    // D[i][j]: lane holds rows 4q..4q+3 of each tile for right-hand side j = lane & 15;
    // part[chunk][v][slot]
    if (vok) {
        double *pp = part + ((size_t)blockIdx.y * k + v) * n_rows;
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int s2 = base + 16 * t + 4 * q + r;
                if (s2 < n_rows) pp[s2] = cd[t][r] + (double)c[t][r];
            }
    }
}

// Row-scaled FP16 operator (A16, the vmult selection and signs included, row scales rs),
// float vectors, on the FP16 matrix cores (v_mfma_f32_16x16x16_f16; layout checked on the
// device: A[i][kk] in lane i + 16 (kk/4), element kk % 4, B likewise with j for i, D as for
// the FP32 instruction).  Each right-hand side is scaled by a power of two xs[v] (its
// largest entry to about 2^13) and split, once per product, into FP16 hi and
// lo = (x - hi) 2^11 (k_vec_split), contracted in two accumulator chains, so the vectors
// keep about 22 bits.  The block stages panels of its 128 rows x HP16 columns through LDS:
// every load instruction of a wavefront reads whole HP16-column row segments (256 bytes at
// 128 columns), where fragment-shaped loads would read 16 rows x 64 bytes.  Lane (row i,
// k-slot q) then takes the 8 contiguous columns cb + 8q .. cb + 8q + 7 of each 32-column
// sub-step from LDS, and the vector halves from the panel's copy in LDS (loaded once per
// block, not once per wavefront: 1.35 -> 1.21 ms at 61k, k = 8); matrix-core pass m of a
// sub-step contracts its columns 8q + 4m + e (e < 4) for A and B alike.  Sums move into FP64 every 256 columns; the row scale and
// 1/xs[v] are applied to the FP64 chunk sums.
typedef _Float16 hvec8 __attribute__((ext_vector_type(8)));
typedef _Float16 hvec4 __attribute__((ext_vector_type(4)));
#ifndef WB_MULTI16_PANEL
#define WB_MULTI16_PANEL 128
#endif
constexpr int HP16 = WB_MULTI16_PANEL;  // columns per staged panel (a multiple of 32)
constexpr int HU16 = HP16 / 8;          // 16-byte units per panel row
constexpr int HS16 = HU16 + 1;          // LDS row stride in units

// xs[v] = 2^(13 - e), e the exponent of max |X[:, v]| (1 for a zero vector); one block
// per vector, fixed reduction order
__global__ __launch_bounds__(1024) void k_vec_scale(const float *__restrict__ X, size_t ldx,
                                                   int n, float *__restrict__ xs)
{
    __shared__ float red[16];
    const float *x = X + (size_t)blockIdx.x * ldx;
    float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, m3 = 0.0f;
    int e = threadIdx.x;
    for (; e + 3 * 1024 < n; e += 4 * 1024) {
        m0 = fmaxf(m0, fabsf(x[e]));
        m1 = fmaxf(m1, fabsf(x[e + 1024]));
        m2 = fmaxf(m2, fabsf(x[e + 2048]));
        m3 = fmaxf(m3, fabsf(x[e + 3072]));
    }
    for (; e < n; e += 1024) m0 = fmaxf(m0, fabsf(x[e]));
    float m = fmaxf(fmaxf(m0, m1), fmaxf(m2, m3));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = 0.0f;
        for (int w = 0; w < 16; ++w) m = fmaxf(m, red[w]);
        int ex = 0;
        if (m > 0.0f) (void)frexpf(m, &ex);
        xs[blockIdx.x] = m > 0.0f ? ldexpf(1.0f, 13 - ex) : 1.0f;
    }
}

// XS[v][c] = fp16(xs[v] X[v][c]) and XS[k + v][c] = fp16((xs[v] X[v][c] - XS[v][c]) 2^11)
// for c < lds (zero past n); one thread per 8 columns
__global__ __launch_bounds__(256) void k_vec_split(const float *__restrict__ X, size_t ldx,
                                                  int n, int k, const float *__restrict__ xs,
                                                  _Float16 *__restrict__ XS, size_t lds)
{
    const int v = blockIdx.y;
    const size_t c = 8 * ((size_t)blockIdx.x * 256 + threadIdx.x);
    if (c >= lds) return;
    const float *x = X + (size_t)v * ldx;
    const float sx = xs[v];
    hvec8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float f = c + e < (size_t)n ? x[c + e] * sx : 0.0f;
        const _Float16 h = (_Float16)f;
        hi[e] = h;
        lo[e] = (_Float16)((f - (float)h) * 2048.0f);
    }
    *(hvec8 *)(XS + (size_t)v * lds + c) = hi;
    *(hvec8 *)(XS + (size_t)(k + v) * lds + c) = lo;
}

__global__ __launch_bounds__(64 * MW) void k_gemm_h16(
    const _Float16 *__restrict__ A16, const float *__restrict__ rs, size_t ld, int n, int k,
    const _Float16 *__restrict__ XS, const float *__restrict__ xs, const int *__restrict__ rows,
    int n_rows, int cols_per_chunk, double *__restrict__ part)
{
    // the block's RPB rows x HP columns of A16, staged in LDS with a row stride of HS
    // 16-byte units (one more than a row holds, so the 16 rows of a fragment read fall in
    // distinct bank groups)
    __shared__ hvec8 tile[RPB * HS16];
    // the panel's columns of the split vectors (hi rows 0 .. k - 1, lo rows k .. 2k - 1),
    // loaded once per block instead of once per wavefront
    __shared__ hvec8 xtile[2 * WB_MULTI_MAX * HS16];
    const int c0 = blockIdx.y * cols_per_chunk, c1 = min(n, c0 + cols_per_chunk);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int brow = blockIdx.x * RPB;
    // staging: thread tid moves unit su of rows sr, sr + SR, ...: a wavefront reads
    // 64 / HU16 whole row segments of HP columns per load instruction
    constexpr int SR = 64 * MW / HU16;   // rows per staging pass
    constexpr int SJ = RPB / SR;         // staging passes
    const int su = tid % HU16, sr = tid / HU16;
    size_t soff[SJ];
#pragma unroll
    for (int j = 0; j < SJ; ++j) soff[j] = (size_t)tile_row(brow + sr + SR * j, rows, n_rows) * ld;
    const int i = lane & 15, q = lane >> 4;  // row in tile, k-slot
    const int v = lane & 15;  // right-hand side fed by this lane's B operand
    const bool vok = v < k;
    fvec4 ch[RT], cl[RT];
    double cd[RT][4];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        ch[t] = fvec4{0.f, 0.f, 0.f, 0.f};
        cl[t] = fvec4{0.f, 0.f, 0.f, 0.f};
        cd[t][0] = cd[t][1] = cd[t][2] = cd[t][3] = 0.0;
    }
    constexpr float ILO = 1.0f / 2048.0f;
    constexpr int FLUSH = 256 / HP16 > 0 ? 256 / HP16 : 1;  // panels per FP64 flush
    int since = 0;
    const hvec8 z8 = {0, 0, 0, 0, 0, 0, 0, 0};
    // A16 is zero past n up to ld; columns of the next chunk read as zero
    hvec8 st[SJ];
    auto load_panel = [&](int pc) {
        const int scol = pc + 8 * su;
#pragma unroll
        for (int j = 0; j < SJ; ++j)
            st[j] = scol < c1 ? __builtin_nontemporal_load((const hvec8 *)(A16 + soff[j] + scol))
                              : z8;
    };
    for (int cb = c0; cb < c1; cb += HP16) {
        if (++since > FLUSH) {
#pragma unroll
            for (int t = 0; t < RT; ++t) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    cd[t][r] += (double)ch[t][r] + (double)cl[t][r] * (double)ILO;
                ch[t] = fvec4{0.f, 0.f, 0.f, 0.f};
                cl[t] = fvec4{0.f, 0.f, 0.f, 0.f};
            }
            since = 1;
        }
        load_panel(cb);
        hvec8 sx[2 * WB_MULTI_MAX * HU16 / (64 * MW) > 0 ? 2 * WB_MULTI_MAX * HU16 / (64 * MW) : 1];
        {
            // unit t of the 2k x HU16 vector panel: vector t / HU16, unit t % HU16
            constexpr int XJ = 2 * WB_MULTI_MAX * HU16 / (64 * MW) > 0
                                   ? 2 * WB_MULTI_MAX * HU16 / (64 * MW)
                                   : 1;
#pragma unroll
            for (int j = 0; j < XJ; ++j) {
                const int t = tid + 64 * MW * j, vv = t / HU16, uu = t % HU16;
                const int col = cb + 8 * uu;
                sx[j] = (vv < 2 * k && col < c1) ? *(const hvec8 *)(XS + (size_t)vv * ld + col)
                                                 : z8;
            }
        }
        __syncthreads();  // the previous panel's fragments have been read
#pragma unroll
        for (int j = 0; j < SJ; ++j) tile[(sr + SR * j) * HS16 + su] = st[j];
        {
            constexpr int XJ = 2 * WB_MULTI_MAX * HU16 / (64 * MW) > 0
                                   ? 2 * WB_MULTI_MAX * HU16 / (64 * MW)
                                   : 1;
#pragma unroll
            for (int j = 0; j < XJ; ++j) {
                const int t = tid + 64 * MW * j;
                if (t < 2 * WB_MULTI_MAX * HU16) xtile[(t / HU16) * HS16 + t % HU16] = sx[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < HP16 / 32; ++u) {
            // this lane's 8 columns of sub-step u: cb + 32 u + 8 q .. + 7 (zero past the
            // chunk in both operands)
            hvec8 bh = z8, bl = z8;
            if (vok) {
                bh = xtile[v * HS16 + 4 * u + q];
                bl = xtile[(k + v) * HS16 + 4 * u + q];
            }
            hvec8 a[RT];
#pragma unroll
            for (int t = 0; t < RT; ++t) a[t] = tile[(wave * 16 * RT + 16 * t + i) * HS16 + 4 * u + q];
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const hvec4 bhm = {bh[4 * m], bh[4 * m + 1], bh[4 * m + 2], bh[4 * m + 3]};
                const hvec4 blm = {bl[4 * m], bl[4 * m + 1], bl[4 * m + 2], bl[4 * m + 3]};
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    const hvec4 am = {a[t][4 * m], a[t][4 * m + 1], a[t][4 * m + 2],
                                      a[t][4 * m + 3]};
                    ch[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(am, bhm, ch[t], 0, 0, 0);
                    cl[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(am, blm, cl[t], 0, 0, 0);
                }
            }
        }
    }
    if (vok) {
        double *pp = part + ((size_t)blockIdx.y * k + v) * n_rows;
        const double isx = 1.0 / (double)xs[v];
        const int base = brow + wave * 16 * RT;
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int s2 = base + 16 * t + 4 * q + r;
                if (s2 < n_rows) {
                    const int rr = rows ? rows[s2] : s2;
                    const double tot = cd[t][r] + (double)ch[t][r] + (double)cl[t][r] * (double)ILO;
                    pp[s2] = tot * (double)rs[rr] * isx;
                }
            }
    }
}

// FP64 accumulation on the FP64 matrix cores: MODE 0 reads the FP64 matrices, MODE 2 the
// FP32 hi + bf16 lo pair (hi + lo formed in FP64 on load); FP64 vectors.  Lane streams 4
// columns (one 4-column group) of its row of each tile.
template <int MODE>
__global__ __launch_bounds__(64 * MW) void k_gemm_sel_d(
    const void *__restrict__ Gv, const void *__restrict__ Hv, const uint16_t *__restrict__ Glo,
    const uint16_t *__restrict__ Hlo, size_t ld, const int4 *__restrict__ segs, int n_segs,
    const uint8_t *__restrict__ gmask, int flip, double s_sel, double s_uns, int k,
    const double *__restrict__ X, size_t ldx, const int *__restrict__ rows, int n_rows,
    int groups_per_chunk, double *__restrict__ part)
{
    const int gc0 = blockIdx.y * groups_per_chunk, gc1 = gc0 + groups_per_chunk;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int base = (blockIdx.x * MW + wave) * 16 * RT;
    const int i = lane & 15, q = lane >> 4;
    size_t roff[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) roff[t] = (size_t)tile_row(base + 16 * t + i, rows, n_rows) * ld;
    const int v = lane & 15;
    const bool vok = v < k;
    const double *Xv = X + (size_t)(vok ? v : 0) * ldx;
    // the four entries of columns col..col+3 of matrix M (0 = G, 1 = H) in row offset ro
    auto load4 = [&](int which, size_t ro, int col, double (&a)[4]) {
        if (MODE == 0) {
            const double *M = (const double *)(which ? Hv : Gv) + ro + col;
            const dvec2 a0 = __builtin_nontemporal_load((const dvec2 *)M);
            const dvec2 a1 = __builtin_nontemporal_load((const dvec2 *)M + 1);
            a[0] = a0.x;
            a[1] = a0.y;
            a[2] = a1.x;
            a[3] = a1.y;
        } else {
            const fvec4 h = __builtin_nontemporal_load(
                (const fvec4 *)((const float *)(which ? Hv : Gv) + ro + col));
            const uvec4 l =
                __builtin_nontemporal_load((const uvec4 *)((which ? Hlo : Glo) + ro + col));
            a[0] = (double)h.x + (double)__uint_as_float((uint32_t)l.x << 16);
            a[1] = (double)h.y + (double)__uint_as_float((uint32_t)l.y << 16);
            a[2] = (double)h.z + (double)__uint_as_float((uint32_t)l.z << 16);
            a[3] = (double)h.w + (double)__uint_as_float((uint32_t)l.w << 16);
        }
    };
    dvec4 c[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) c[t] = dvec4{0.0, 0.0, 0.0, 0.0};
    for (int si = 0; si < n_segs; ++si) {
        const int4 sg = segs[si];
        const int g_lo = max(sg.x, gc0), g_hi = min(sg.y, gc1);  // this chunk's part
        if (g_lo >= g_hi) continue;
        const int c_end = 4 * g_hi;
        const unsigned fm = flip ? 0xfu : 0u;
        const bool pure = sg.z < 2;
        const int use_h = (sg.z ^ flip) != 0;
        for (int cb = 4 * g_lo; cb < c_end; cb += 16) {
            const int col = cb + 4 * q;  // this lane's 4 columns (one 4-column group)
            double a[RT][4], b[4] = {0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < RT; ++t) a[t][0] = a[t][1] = a[t][2] = a[t][3] = 0.0;
            if (col < c_end) {
                double sgn[4];
                if (pure) {
#pragma unroll
                    for (int t = 0; t < RT; ++t) load4(use_h, roff[t], col, a[t]);
                    const double s = use_h ? s_sel : s_uns;
                    sgn[0] = sgn[1] = sgn[2] = sgn[3] = s;
                } else {
                    const unsigned m = gmask[col >> 2] ^ fm;
#pragma unroll
                    for (int t = 0; t < RT; ++t) {
                        double ag[4], ah[4];
                        load4(0, roff[t], col, ag);
                        load4(1, roff[t], col, ah);
#pragma unroll
                        for (int e = 0; e < 4; ++e) a[t][e] = ((m >> e) & 1) ? ah[e] : ag[e];
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) sgn[e] = ((m >> e) & 1) ? s_sel : s_uns;
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
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int t = 0; t < RT; ++t)
                    c[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t][s], b[s], c[t], 0, 0, 0);
        }
    }
    // FP64 result layout: D[i][j] in lane j + 16 (i % 4), register i / 4
    if (vok) {
        double *pp = part + ((size_t)blockIdx.y * k + v) * n_rows;
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int s2 = base + 16 * t + q + 4 * r;
                if (s2 < n_rows) pp[s2] = c[t][r];
            }
    }
}

// The hi + bf16 lo operator (FP64 vectors) with LDS-staged panels: the block's 128 rows x
// HL_P columns of the selected hi (FP32) and lo (bf16) entries are staged with whole-row
// loads (at 64 columns a wavefront reads 4 rows x 256 bytes of hi and 4 rows x 128 bytes of
// lo per instruction), the panel's vector entries once per block with their signs, and
// contracted on the FP64 matrix cores as in k_gemm_sel_d.  A staging
// thread owns one 4-column group of the panel: gmask says whether its columns read G or H
// (both, column by column, for a mixed group); the signs go to the vector operand.
// LDS row strides of HL_U + 1 (hi, 16-byte units) and HL_U + 2 (lo, 8-byte units) put the
// 16 rows of a fragment read in distinct banks.  At 61k, k = 8: 64-column panels, 2 blocks
// per CU (61.5 KB of LDS each), 4.68 ms; 32 columns, 4 blocks per CU, 4.96 ms; 16 columns
// 8.87 ms.
#ifndef WB_MULTI_HL_PANEL
#define WB_MULTI_HL_PANEL 64
#endif
constexpr int HL_P = WB_MULTI_HL_PANEL;  // columns per panel (a multiple of 16)
constexpr int HL_U = HL_P / 4;      // 4-column groups per panel row
constexpr int HL_SH = HL_U + 1;     // hi row stride (16-byte units)
constexpr int HL_SL = HL_U + 2;     // lo row stride (8-byte units)

#ifndef WB_MULTI_HL_WAVES
#define WB_MULTI_HL_WAVES 2  // wavefronts per SIMD the hi + lo kernel is built for (the
                             // blocks per CU its LDS allows)
#endif

__global__ __launch_bounds__(64 * MW, WB_MULTI_HL_WAVES) void k_gemm_hl_lds(
    const float *__restrict__ G, const float *__restrict__ H, const uint16_t *__restrict__ Glo,
    const uint16_t *__restrict__ Hlo, size_t ld, int n, const uint8_t *__restrict__ gmask,
    int flip, double s_sel, double s_uns, int k, const double *__restrict__ X, size_t ldx,
    const int *__restrict__ rows, int n_rows, int cols_per_chunk, double *__restrict__ part)
{
    __shared__ fvec4 tile_h[RPB * HL_SH];
    __shared__ uvec4 tile_l[RPB * HL_SL];
    // the panel's vector entries, signs applied: [v][column] (loaded once per block)
    __shared__ double xtile[WB_MULTI_MAX * (HL_P + 1)];
    const int c0 = blockIdx.y * cols_per_chunk, c1 = min(n, c0 + cols_per_chunk);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int brow = blockIdx.x * RPB;
    constexpr int SR = 64 * MW / HL_U;  // rows per staging pass
    constexpr int SJ = RPB / SR;        // staging passes
    const int su = tid % HL_U, sr = tid / HL_U;
    // staged rows as 32-bit indices (their 64-bit offsets are formed at each load, which
    // keeps 4 registers free)
    int srow[SJ];
#pragma unroll
    for (int j = 0; j < SJ; ++j) srow[j] = tile_row(brow + sr + SR * j, rows, n_rows);
    const int i = lane & 15, q = lane >> 4;
    const int v = lane & 15;
    const bool vok = v < k;
    const unsigned fm = flip ? 0xfu : 0u;
    dvec4 c[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) c[t] = dvec4{0.0, 0.0, 0.0, 0.0};
    fvec4 sh[SJ];
    uvec4 sl[SJ];
    // stage: group su of the panel at pc for rows sr + SR j
    auto load_panel = [&](int pc) {
        // an empty asm that "changes" the rows keeps their 64-bit offsets from being
        // hoisted out of the panel loop into registers
#pragma unroll
        for (int j = 0; j < SJ; ++j) asm volatile("" : "+v"(srow[j]));
        const int scol = pc + 4 * su;
        const fvec4 z4 = {0.f, 0.f, 0.f, 0.f};
        const uvec4 zl = {0, 0, 0, 0};
        if (scol < c1) {
            const unsigned m = (unsigned)gmask[scol >> 2] ^ fm;  // bit e: column e reads H
            if (m == 0u || m == 0xfu) {
                const float *M = m ? H : G;
                const uint16_t *L = m ? Hlo : Glo;
#pragma unroll
                for (int j = 0; j < SJ; ++j) {
                    sh[j] = __builtin_nontemporal_load((const fvec4 *)(M + (size_t)srow[j] * ld + scol));
                    sl[j] = __builtin_nontemporal_load((const uvec4 *)(L + (size_t)srow[j] * ld + scol));
                }
            } else {
#pragma unroll
                for (int j = 0; j < SJ; ++j) {
                    const fvec4 g = *(const fvec4 *)(G + (size_t)srow[j] * ld + scol);
                    const fvec4 h = *(const fvec4 *)(H + (size_t)srow[j] * ld + scol);
                    const uvec4 gl = *(const uvec4 *)(Glo + (size_t)srow[j] * ld + scol);
                    const uvec4 hl = *(const uvec4 *)(Hlo + (size_t)srow[j] * ld + scol);
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const bool he = (m >> e) & 1u;
                        sh[j][e] = he ? h[e] : g[e];
                        sl[j][e] = he ? hl[e] : gl[e];
                    }
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < SJ; ++j) {
                sh[j] = z4;
                sl[j] = zl;
            }
        }
    };
    for (int cb = c0; cb < c1; cb += HL_P) {
        load_panel(cb);
        constexpr int XE = (WB_MULTI_MAX * HL_P + 64 * MW - 1) / (64 * MW);
        double sx[XE];
#pragma unroll
        for (int j = 0; j < XE; ++j) {
            const int t = tid + 64 * MW * j, vv = t / HL_P, cc = cb + t % HL_P;
            double xv = 0.0;
            if (vv < k && cc < c1) {
                const unsigned m = (unsigned)gmask[cc >> 2] ^ fm;
                xv = X[(size_t)vv * ldx + cc] * (((m >> (cc & 3)) & 1u) ? s_sel : s_uns);
            }
            sx[j] = xv;
        }
        __syncthreads();  // the previous panel's fragments have been read
#pragma unroll
        for (int j = 0; j < SJ; ++j) {
            tile_h[(sr + SR * j) * HL_SH + su] = sh[j];
            tile_l[(sr + SR * j) * HL_SL + su] = sl[j];
        }
#pragma unroll
        for (int j = 0; j < XE; ++j) {
            const int t = tid + 64 * MW * j;
            if (t < WB_MULTI_MAX * HL_P) xtile[(t / HL_P) * (HL_P + 1) + t % HL_P] = sx[j];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < HL_P / 16; ++u) {
            // this lane's 4 columns cb + 16 u + 4 q .. + 3, signs applied (zero past the chunk)
            double b[4] = {0.0, 0.0, 0.0, 0.0};
            if (vok) {
#pragma unroll
                for (int e = 0; e < 4; ++e) b[e] = xtile[v * (HL_P + 1) + 16 * u + 4 * q + e];
            }
            double a[RT][4];
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                const int r = wave * 16 * RT + 16 * t + i;
                const fvec4 h = tile_h[r * HL_SH + 4 * u + q];
                const uvec4 l = tile_l[r * HL_SL + 4 * u + q];
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    a[t][e] = (double)h[e] + (double)__uint_as_float((uint32_t)l[e] << 16);
            }
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
                for (int t = 0; t < RT; ++t)
                    c[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t][s2], b[s2], c[t], 0, 0, 0);
        }
    }
    // FP64 result layout: D[i][j] in lane j + 16 (i % 4), register i / 4
    if (vok) {
        double *pp = part + ((size_t)blockIdx.y * k + v) * n_rows;
        const int base = brow + wave * 16 * RT;
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int s2 = base + 16 * t + q + 4 * r;
                if (s2 < n_rows) pp[s2] = c[t][r];
            }
    }
}

// Y[v][rows[s]] = sum over column chunks of part[c][v][s] (in order) + diag .* Xd
template <typename T>
__global__ void k_multi_reduce(const double *__restrict__ part, int n_chunks, int k,
                               const int *__restrict__ rows, int n_rows,
                               const double *__restrict__ diag, const T *__restrict__ Xd,
                               size_t ldx, T *__restrict__ Y, size_t ldy)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x, v = blockIdx.y;
    if (s >= n_rows || v >= k) return;
    const int rr = rows ? rows[s] : s;
    double val = 0.0;
    for (int c = 0; c < n_chunks; ++c) val += part[((size_t)c * k + v) * n_rows + s];
    if (diag) val += diag[rr] * (double)Xd[(size_t)v * ldx + rr];
    Y[(size_t)v * ldy + rr] = (T)val;
}

}  // namespace

// Column chunks of the batched product.  The matrix-core kernels of modes 0-2 take about
// 8k wavefronts whatever n is.  The FP16 kernel (mode 3) streams the operator at close to
// the memory rate, so its grid is sized in whole rounds of resident blocks: the chunk count
// (1..16) whose block count comes closest to a multiple of what the device holds at once,
// the smaller count on ties.
int multi_chunks(int mode, int n_rows, int n_groups)
{
    // the resident-block count is cached (a failed query stays 0 and selects the rule of the
    // other modes below: a grid shape only, the sums are the same), so the workspace sizing
    // and the launch always agree
    const int cap = mode == 3 ? resident_blocks(reinterpret_cast<const void *>(&k_gemm_h16),
                                                64 * MW)
                              : 0;
    if (mode == 3 && cap > 0) {
        const long long blocks = (n_rows + RPB - 1) / RPB;
        const int steps = (4 * n_groups + HP16 - 1) / HP16;
        int best = 1;
        double best_fill = 0.0;
        for (int c = 1; c <= 16 && c <= std::max(1, steps); ++c) {
            const long long b = blocks * c, rounds = (b + cap - 1) / cap;
            const double fill = (double)b / (double)(rounds * cap);
            if (fill > best_fill + 1e-3) {
                best_fill = fill;
                best = c;
            }
        }
        return best;
    }
    const int waves = (n_rows + 16 * RT - 1) / (16 * RT);
    int c = (8192 + waves - 1) / waves;
    c = c < 1 ? 1 : (c > 16 ? 16 : c);
    return c > n_groups ? (n_groups > 0 ? n_groups : 1) : c;
}

void launch_gemv_sel_multi(int mode, const void *G, const void *H, const void *Glo,
                           const void *Hlo, size_t ld, const int *segs, int n_segs,
                           const uint8_t *gmask, int flip, double s_sel, double s_unsel, int k,
                           const void *X, size_t ldx, const double *diag, const void *Xd,
                           const int *rows, int n_rows, void *Y, size_t ldy, double *part,
                           int n_groups, hipStream_t s)
{
    if (n_rows <= 0 || k <= 0) return;
    const int chunks = multi_chunks(mode, n_rows, n_groups);
    const int gpc = (n_groups + chunks - 1) / chunks;
    dim3 grid((n_rows + RPB - 1) / RPB, chunks), block(64 * MW);
    if (mode == 3) {
        // G holds A16 and H the row scales; Glo is the workspace: k vector scales, then the
        // split vectors (2k x ld halfs from byte 256); chunks of whole HP16-column panels
        const int n = 4 * n_groups;
        constexpr int step = HP16;
        const int cpc = ((n + chunks - 1) / chunks + step - 1) / step * step;
        float *xs = (float *)const_cast<void *>(Glo);
        _Float16 *XS = (_Float16 *)((char *)xs + 256);
        hipLaunchKernelGGL(k_vec_scale, dim3(k), dim3(1024), 0, s, (const float *)X, ldx, n, xs);
        hipLaunchKernelGGL(k_vec_split, dim3((unsigned)((ld / 8 + 255) / 256), k), dim3(256), 0,
                           s, (const float *)X, ldx, n, k, (const float *)xs, XS, ld);
        hipLaunchKernelGGL(k_gemm_h16, grid, block, 0, s, (const _Float16 *)G, (const float *)H,
                           ld, n, k, (const _Float16 *)XS, (const float *)xs, rows, n_rows, cpc,
                           part);
    } else if (mode == 1)
        hipLaunchKernelGGL(k_gemm_sel_f32, grid, block, 0, s, (const float *)G, (const float *)H,
                           ld, (const int4 *)segs, n_segs, gmask, flip, (float)s_sel,
                           (float)s_unsel, k, (const float *)X, ldx, rows, n_rows, gpc, part);
    else if (mode == 2) {
        // chunks of whole HL_P-column panels (chunks past the columns write zero sums)
        const int n = 4 * n_groups;
        const int cpc = ((n + chunks - 1) / chunks + HL_P - 1) / HL_P * HL_P;
        hipLaunchKernelGGL(k_gemm_hl_lds, grid, block, 0, s, (const float *)G, (const float *)H,
                           (const uint16_t *)Glo, (const uint16_t *)Hlo, ld, n, gmask, flip,
                           s_sel, s_unsel, k, (const double *)X, ldx, rows, n_rows, cpc, part);
    }
    else
        hipLaunchKernelGGL(k_gemm_sel_d<0>, grid, block, 0, s, G, H, nullptr, nullptr, ld,
                           (const int4 *)segs, n_segs, gmask, flip, s_sel, s_unsel, k,
                           (const double *)X, ldx, rows, n_rows, gpc, part);
    dim3 rg((n_rows + 255) / 256, k);
    if (mode == 1 || mode == 3)
        hipLaunchKernelGGL(k_multi_reduce<float>, rg, dim3(256), 0, s, part, chunks, k, rows,
                           n_rows, diag, (const float *)Xd, ldx, (float *)Y, ldy);
    else
        hipLaunchKernelGGL(k_multi_reduce<double>, rg, dim3(256), 0, s, part, chunks, k, rows,
                           n_rows, diag, (const double *)Xd, ldx, (double *)Y, ldy);
}

}  // namespace wb
