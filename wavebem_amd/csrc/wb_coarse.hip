// wb_coarse.hip -- coarse-space correction of the block-Jacobi preconditioner on gfx950.
//
// Two-level (multiplicative) preconditioner for the GMRES of BEMProblem::solve_system
// (source/bem_problem.cc:821-895; the reference preconditions with a banded LU,
// :1107-1149, and any preconditioner leaves the converged solution unchanged):
//   M^-1 v = P z + B^-1 (v - A P z),   z = Ac^-1 P^T v,   Ac = P^T A P,
// with B the block-Jacobi blocks and P the indicator of aggregates: runs of `agg`
// consecutive dofs of the device ordering (compact patches, since the ordering is built
// from compact column tiles), agg a multiple of 8 and at least 64 (so a block-Jacobi block
// of at most 128 rows touches at most 3 aggregates), at most CS_MAX aggregates so that Ac
// is inverted in the LDS of one workgroup.  A is the matrix the block-Jacobi blocks are
// taken from: H(j,i) (+ alpha_i at i == j) in a Neumann column, -G(j,i) in a Dirichlet
// column, identity rows for constrained dofs.  Block Jacobi handles the near field; the
// coarse space takes the smooth coupling between patches, which it misses.
//
// Per build: AP = A P (one pass over the selected operator: the FP64 matrices; in FP32-IR
// its row-scaled FP16 copy, or the FP32 hi part with the FP32 inner format), Ac = P^T AP,
// its inverse,
// and W = AP Ac^-1 (n x nc).  Per application: C = P^T v (one block per aggregate),
// t = v - W C, then the block-Jacobi kernel applies B^-1 t and adds z = Ac^-1 C of the
// row's aggregate.  The transposed system (the adjoint solve) uses the same aggregates:
// its coarse matrix is Ac^T, and W = A^T P Ac^-T needs one more pass over A for P^T A.
// Every sum runs in a fixed order, so results do not depend on timing.
#include "wb_device.h"
#include "wb_internal.h"

namespace wb {
namespace {

__device__ __forceinline__ double cs_wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

constexpr int CR_THREADS = 1024;  // k_coarse_rows: one row per block
constexpr int CI_THREADS = 1024;  // k_coarse_invert

typedef float cs_fvec4 __attribute__((ext_vector_type(4)));
typedef double cs_dvec2 __attribute__((ext_vector_type(2)));

typedef _Float16 cs_hvec8 __attribute__((ext_vector_type(8)));

// AP[a][j] = sum over the columns i of aggregate a of A(j, i): one row j per block, tpa
// consecutive threads per aggregate (a power of two with nc tpa <= CR_THREADS), each
// summing every tpa-th four-column group of the aggregate (gmask bit e of group g: column
// 4g + e comes from H) or, given the row-scaled FP16 copy of the selected operator (A16,
// row scales rs16), every tpa-th eight-column group of it; log2(tpa) shuffle steps join
// them.  Padding columns past n are zero in G, H and A16.
__global__ __launch_bounds__(CR_THREADS) void k_coarse_rows(
    int n, int agg, int nc, const void *__restrict__ Gv, const void *__restrict__ Hv,
    int precision, size_t ld, const _Float16 *__restrict__ A16, const float *__restrict__ rs16,
    const double *__restrict__ alpha, const uint8_t *__restrict__ gmask,
    const uint8_t *__restrict__ constrained, double *__restrict__ AP, size_t ldap)
{
    const int j = blockIdx.x;
    const bool con = constrained[j] != 0;  // identity row
    const size_t roff = (size_t)j * ld;
    int tpa = 1;
    while (tpa < 64 && 2 * tpa * nc <= CR_THREADS) tpa *= 2;
    const int a = (int)threadIdx.x / tpa, q = (int)threadIdx.x % tpa;
    const int c0 = a * agg, c1 = min(n, c0 + agg);
    double s = 0.0;
    if (a < nc && !con) {
        if (A16) {
            const int h1 = (c1 + 7) / 8;  // agg is a multiple of 8
            for (int h = c0 / 8 + q; h < h1; h += tpa) {
                const cs_hvec8 v = *((const cs_hvec8 *)(A16 + roff) + h);
                s += (double)((((float)v[0] + (float)v[1]) + ((float)v[2] + (float)v[3])) +
                              (((float)v[4] + (float)v[5]) + ((float)v[6] + (float)v[7])));
            }
        } else {
            const int g1 = (c1 + 3) / 4;
            for (int g = c0 / 4 + q; g < g1; g += tpa) {
                const unsigned m = gmask[g];
                double e[4];
                if (precision == WB_PREC_FP64) {
                    const double *Gr = (const double *)Gv + roff + 4 * (size_t)g;
                    const double *Hr = (const double *)Hv + roff + 4 * (size_t)g;
                    cs_dvec2 g01 = {0, 0}, g23 = {0, 0}, h01 = {0, 0}, h23 = {0, 0};
                    if (m != 15u) {
                        g01 = *(const cs_dvec2 *)Gr;
                        g23 = *((const cs_dvec2 *)Gr + 1);
                    }
                    if (m != 0u) {
                        h01 = *(const cs_dvec2 *)Hr;
                        h23 = *((const cs_dvec2 *)Hr + 1);
                    }
                    e[0] = (m & 1u) ? h01.x : -g01.x;
                    e[1] = (m & 2u) ? h01.y : -g01.y;
                    e[2] = (m & 4u) ? h23.x : -g23.x;
                    e[3] = (m & 8u) ? h23.y : -g23.y;
                } else {
                    // the FP32 hi part is accurate enough for a preconditioner
                    cs_fvec4 gv = {0, 0, 0, 0}, hv = {0, 0, 0, 0};
                    if (m != 15u) gv = *((const cs_fvec4 *)((const float *)Gv + roff) + g);
                    if (m != 0u) hv = *((const cs_fvec4 *)((const float *)Hv + roff) + g);
                    e[0] = (m & 1u) ? hv.x : -gv.x;
                    e[1] = (m & 2u) ? hv.y : -gv.y;
                    e[2] = (m & 4u) ? hv.z : -gv.z;
                    e[3] = (m & 8u) ? hv.w : -gv.w;
                }
                s += (e[0] + e[1]) + (e[2] + e[3]);
            }
        }
    }
    for (int o = 1; o < tpa; o <<= 1) s += __shfl_xor(s, o, 64);
    if (a < nc && q == 0) {
        const bool mine = j >= c0 && j < c1;
        if (con) {
            s = mine ? 1.0 : 0.0;
        } else {
            if (A16) s *= (double)rs16[j];
            // the alpha diagonal of a Neumann column
            if (mine && ((gmask[j >> 2] >> (j & 3)) & 1u)) s += alpha[j];
        }
        AP[(size_t)a * ldap + j] = s;
    }
}

// The same rows of A P from the aggregate sums S of the FP16 copy (written by its build):
// AP[a][j] = S[a][j] plus the alpha diagonal of a Neumann column; identity rows for
// constrained dofs.  One thread per (j, a).
__global__ __launch_bounds__(256) void k_coarse_rows_sums(
    int n, int agg, int nc, const double *__restrict__ S, const double *__restrict__ alpha,
    const uint8_t *__restrict__ gmask, const uint8_t *__restrict__ constrained,
    double *__restrict__ AP, size_t ldap)
{
    const int j = blockIdx.x * 256 + threadIdx.x, a = blockIdx.y;
    if (j >= n) return;
    const bool mine = j >= a * agg && j < min(n, (a + 1) * agg);
    const size_t o = (size_t)a * ldap + j;
    double s;
    if (constrained[j] != 0)
        s = mine ? 1.0 : 0.0;
    else
        s = S[o] + (mine && ((gmask[j >> 2] >> (j & 3)) & 1u) ? alpha[j] : 0.0);
    AP[o] = s;
}

// Ac[a][b] = sum over rows j of aggregate a of AP[b][j]: one wavefront per (a, b).
__global__ __launch_bounds__(256) void k_coarse_matrix(int n, int agg, int nc,
                                                       const double *__restrict__ AP, size_t ldap,
                                                       double *__restrict__ Ac)
{
    const int lane = threadIdx.x & 63;
    const int pair = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pair >= nc * nc) return;  // wave-uniform
    const int a = pair / nc, b = pair % nc;
    const int r0 = a * agg, r1 = min(n, r0 + agg);
    double s = 0.0;
    for (int rb = r0; rb < r1; rb += 64) {
        const int j = rb + lane;
        s += cs_wave_sum(j < r1 ? AP[(size_t)b * ldap + j] : 0.0);
    }
    if (lane == 0) Ac[a * nc + b] = s;
}

// Aci = Ac^-1 (row-major, FP64 and FP32), eliminated in LDS in type T: FP64 up to 128
// unknowns, FP32 above (to CS_MAX = 192; the inverse only preconditions, and the FP64 path
// applies the stored values as a fixed linear operator); status[0] = zero pivots + 1 if
// any entry of the inverse is not finite (0: usable).
template <typename T>
__global__ __launch_bounds__(CI_THREADS) void k_coarse_invert(int nc, const double *__restrict__ Ac,
                                                              double *__restrict__ Aci,
                                                              float *__restrict__ Aci32,
                                                              int *__restrict__ status)
{
    extern __shared__ double sm[];
    T *A = (T *)sm;                                      // nc x nc
    T *colk = A + (size_t)nc * nc;                       // nc
    int *piv = (int *)(colk + nc);                       // nc
    __shared__ int red_i[1];
    for (int e = threadIdx.x; e < nc * nc; e += CI_THREADS) A[e] = (T)Ac[e];
    __syncthreads();
    const int zp = gj_invert_lds<CI_THREADS, T>(A, colk, piv, red_i, nc);
    int bad = 0;
    for (int e = threadIdx.x; e < nc * nc; e += CI_THREADS) {
        const double v = (double)A[e];
        bad |= !isfinite(v);
        Aci[e] = v;
        Aci32[e] = (float)v;
    }
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) status[0] = zp + (bad ? 1 : 0);
}

// WT[b][i] = (AP Ac^-1)[i][b] = sum_a AP[a][i] Aci[a][b] (with trans: Aci[b][a], i.e.
// Ac^-T, for the transposed system): thread per row i, CW_TILE columns b per block
// (blockIdx.y), so AP is read once per CW_TILE columns; WT32 may be null.  At 41k rows and
// 128 columns, 8 columns per block took 0.25 ms, 16 took 0.30 ms and 32 took 0.34 ms.
#ifndef WB_CW_TILE
#define WB_CW_TILE 8
#endif
constexpr int CW_TILE = WB_CW_TILE;
__global__ __launch_bounds__(256) void k_coarse_w(int n, int nc, const double *__restrict__ AP,
                                                  size_t ldap, const double *__restrict__ Aci,
                                                  int trans, double *__restrict__ WT,
                                                  float *__restrict__ WT32)
{
    const int b0 = blockIdx.y * CW_TILE;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc[CW_TILE];
#pragma unroll
    for (int t = 0; t < CW_TILE; ++t) acc[t] = 0.0;
    for (int a = 0; a < nc; ++a) {
        const double ap = AP[(size_t)a * ldap + i];
#pragma unroll
        for (int t = 0; t < CW_TILE; ++t)
            if (b0 + t < nc)
                acc[t] += ap * (trans ? Aci[(size_t)(b0 + t) * nc + a] : Aci[(size_t)a * nc + b0 + t]);
    }
#pragma unroll
    for (int t = 0; t < CW_TILE; ++t)
        if (b0 + t < nc) {
            WT[(size_t)(b0 + t) * ldap + i] = acc[t];
            if (WT32) WT32[(size_t)(b0 + t) * ldap + i] = (float)acc[t];
        }
}

// For the transposed system: PTA[a][i] = sum over the rows j of aggregate a of A(j, i),
// i.e. row i of A^T P (FP64 matrices): thread per column i, rows in order.
__global__ __launch_bounds__(256) void k_coarse_cols(
    int n, int agg, const double *__restrict__ G, const double *__restrict__ H, size_t ld,
    const double *__restrict__ alpha, const uint8_t *__restrict__ gmask,
    const uint8_t *__restrict__ constrained, double *__restrict__ PTA, size_t ldap)
{
    const int a = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool neu = (gmask[i >> 2] >> (i & 3)) & 1u;
    const double *M = neu ? H : G;
    const int r0 = a * agg, r1 = min(n, r0 + agg);
    double s = 0.0;
    for (int j = r0; j < r1; ++j) {
        double v;
        if (constrained[j])
            v = (i == j) ? 1.0 : 0.0;
        else {
            const double m = M[(size_t)j * ld + i];
            v = neu ? m + (i == j ? alpha[i] : 0.0) : -m;
        }
        s += v;
    }
    PTA[(size_t)a * ldap + i] = s;
}

// C[v][a] = (P^T x_v)[a]: one block per (aggregate, vector); per-thread strided sums,
// then the 4 wavefront sums in order
template <typename T>
__global__ __launch_bounds__(256) void k_coarse_restrict(int n, int agg, const T *__restrict__ X,
                                                         size_t ldx, double *__restrict__ C)
{
    __shared__ double ws[4];
    const int a = blockIdx.x, v = blockIdx.y;
    const T *x = X + (size_t)v * ldx;
    const int r0 = a * agg, r1 = min(n, r0 + agg);
    double s = 0.0;
    for (int i = r0 + (int)threadIdx.x; i < r1; i += 256) s += (double)x[i];
    s = cs_wave_sum(s);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) C[(size_t)v * CS_MAX + a] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

// t_v = x_v - W C[v] (W stored [b][i]); vector v = blockIdx.y.  64 rows per block: the
// aggregates split into four contiguous quarters, each summed by one thread of the row
// with two alternating accumulators, the parts then added in a fixed order, so the chains
// are short and the sum is the same on every call.
template <typename T>
__global__ __launch_bounds__(256) void k_coarse_residual(int n, int nc, const T *__restrict__ X,
                                                         size_t ldx, const T *__restrict__ WT,
                                                         size_t ldw, const double *__restrict__ C,
                                                         T *__restrict__ Tout)
{
    __shared__ double cs[CS_MAX];
    __shared__ double qs[4][64];
    const int v = blockIdx.y;
    for (int b = threadIdx.x; b < nc; b += blockDim.x) cs[b] = C[(size_t)v * CS_MAX + b];
    __syncthreads();
    const int rl = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + rl;
    const int b0 = q * nc / 4, b1 = (q + 1) * nc / 4;
    double a0 = 0.0, a1 = 0.0;
    if (i < n) {
        int b = b0;
        for (; b + 1 < b1; b += 2) {
            a0 = fma((double)WT[(size_t)b * ldw + i], cs[b], a0);
            a1 = fma((double)WT[(size_t)(b + 1) * ldw + i], cs[b + 1], a1);
        }
        if (b < b1) a0 = fma((double)WT[(size_t)b * ldw + i], cs[b], a0);
    }
    qs[q][rl] = a0 + a1;
    __syncthreads();
    if (q == 0 && i < n)
        Tout[(size_t)v * ldx + i] =
            (T)((double)X[(size_t)v * ldx + i] - (((qs[0][rl] + qs[1][rl]) + qs[2][rl]) + qs[3][rl]));
}

// The block-Jacobi apply of one vector with the coarse residual formed in the same
// workgroup (blocks of bs <= 64 rows, 256 threads): t = x - W C for the block's rows, four
// quarter sums per row as in k_coarse_residual; z_a = (Aci C)_a for the aggregates the rows
// touch, one wavefront each; then y = B^-1 t + z, the columns of B^-1 split over the four
// wavefronts and the partial products added in wavefront order.  x and y may alias: every
// row is read and written by its own workgroup only, the read before the write.
#ifndef WB_BJC_WAVES
#define WB_BJC_WAVES 4
#endif
constexpr int BJC_W = WB_BJC_WAVES;  // wavefronts per 64-row block

template <typename T>
__global__ __launch_bounds__(64 * BJC_W) void k_bj_coarse(int n, int bs, const T *__restrict__ binv,
                                                   const T *x, T *y, const T *__restrict__ WT,
                                                   size_t ldw, const double *__restrict__ C,
                                                   const T *__restrict__ Aci, int nc, int agg,
                                                   const double *__restrict__ Cpart)
{
    __shared__ double cs[CS_MAX];
    __shared__ double qs[BJC_W][64];
    __shared__ T ts[64];
    __shared__ T ps[BJC_W][64];
    __shared__ double zs[4];
    const int tid = threadIdx.x, rl = tid & 63, w = tid >> 6;
    const int base = blockIdx.x * bs, m = min(bs, n - base);
    for (int b = tid; b < nc; b += 64 * BJC_W) {
        double v;
        if (Cpart) {
            // aggregate b from the partial sums of the 64-row blocks it overlaps, in order
            const int r0 = b * agg, r1 = min(n, (b + 1) * agg);
            v = 0.0;
            for (int blk = r0 / 64; blk <= (r1 - 1) / 64; ++blk)
                v += Cpart[2 * blk + (b - (blk * 64) / agg)];
        } else {
            v = C[b];
        }
        cs[b] = v;
    }
    __syncthreads();
    const int i = base + rl;
    const int b0 = w * nc / BJC_W, b1 = (w + 1) * nc / BJC_W;
    // loads in batches of 16 (independent, so one wait per batch), two accumulators
    double a0 = 0.0, a1 = 0.0;
    if (rl < m) {
        int b = b0;
        for (; b + 16 <= b1; b += 16) {
            T wv[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) wv[u] = WT[(size_t)(b + u) * ldw + i];
#pragma unroll
            for (int u = 0; u < 16; u += 2) {
                a0 = fma((double)wv[u], cs[b + u], a0);
                a1 = fma((double)wv[u + 1], cs[b + u + 1], a1);
            }
        }
        for (; b < b1; ++b) a0 = fma((double)WT[(size_t)b * ldw + i], cs[b], a0);
    }
    qs[w][rl] = a0 + a1;
    const int ag0 = base / agg, na = (base + m - 1) / agg - ag0 + 1;
    if (w < na) {
        const int a = ag0 + w;
        double p = 0.0;
#pragma unroll 3
        for (int bb = rl; bb < nc; bb += 64) p += (double)Aci[(size_t)a * nc + bb] * cs[bb];
        p = cs_wave_sum(p);
        if (rl == 0) zs[w] = p;
    }
    __syncthreads();
    if (w == 0) {
        double q = qs[0][rl];
#pragma unroll
        for (int u = 1; u < BJC_W; ++u) q += qs[u][rl];
        ts[rl] = rl < m ? (T)((double)x[i] - q) : (T)0;
    }
    __syncthreads();
    const int c0 = w * bs / BJC_W, c1 = min(m, (w + 1) * bs / BJC_W);
    const T *B = binv + (size_t)blockIdx.x * bs * bs;
    T acc = 0;
    if (rl < m) {
        int c = c0;
        constexpr int CB = 64 / BJC_W;  // columns per wavefront at bs = 64, loaded together
        for (; c + CB <= c1; c += CB) {
            T bv[CB];
#pragma unroll
            for (int u = 0; u < CB; ++u) bv[u] = B[(size_t)(c + u) * bs + rl];
#pragma unroll
            for (int u = 0; u < CB; ++u) acc += bv[u] * ts[c + u];
        }
        for (; c < c1; ++c) acc += B[(size_t)c * bs + rl] * ts[c];
    }
    ps[w][rl] = acc;
    __syncthreads();
    if (w == 0 && rl < m) {
        T p = ps[0][rl];
#pragma unroll
        for (int u = 1; u < BJC_W; ++u) p += ps[u][rl];
        y[i] = (T)((double)p + zs[i / agg - ag0]);
    }
}

}  // namespace

size_t coarse_invert_lds_bytes(int nc)
{
    const size_t es = nc <= CS_MAX64 ? sizeof(double) : sizeof(float);
    return es * ((size_t)nc * nc + nc) + sizeof(int) * (size_t)nc;
}

void launch_coarse_build(int n, int agg, int nc, const void *G, const void *H, int precision,
                         size_t ld, const void *A16, const float *rs16, const double *alpha,
                         const uint8_t *gmask, const uint8_t *constrained, double *AP,
                         size_t ldap, const double *S, double *Ac, double *Aci,
                         float *Aci32, double *WT, float *WT32, int *status, hipStream_t s)
{
    if (S)
        hipLaunchKernelGGL(k_coarse_rows_sums, dim3((n + 255) / 256, nc), dim3(256), 0, s, n,
                           agg, nc, S, alpha, gmask, constrained, AP, ldap);
    else
        hipLaunchKernelGGL(k_coarse_rows, dim3(n), dim3(CR_THREADS), 0, s, n, agg, nc, G, H,
                           precision, ld, (const _Float16 *)A16, rs16, alpha, gmask,
                           constrained, AP, ldap);
    hipLaunchKernelGGL(k_coarse_matrix, dim3((nc * nc + 3) / 4), dim3(256), 0, s, n, agg, nc,
                       (const double *)AP, ldap, Ac);
    if (nc <= CS_MAX64)
        hipLaunchKernelGGL(k_coarse_invert<double>, dim3(1), dim3(CI_THREADS),
                           coarse_invert_lds_bytes(nc), s, nc, (const double *)Ac, Aci, Aci32,
                           status);
    else
        hipLaunchKernelGGL(k_coarse_invert<float>, dim3(1), dim3(CI_THREADS),
                           coarse_invert_lds_bytes(nc), s, nc, (const double *)Ac, Aci, Aci32,
                           status);
    hipLaunchKernelGGL(k_coarse_w, dim3((n + 255) / 256, (nc + CW_TILE - 1) / CW_TILE), dim3(256),
                       0, s, n, nc, (const double *)AP, ldap, (const double *)Aci, 0, WT, WT32);
}

void launch_coarse_build_t(int n, int agg, int nc, const double *G, const double *H, size_t ld,
                           const double *alpha, const uint8_t *gmask, const uint8_t *constrained,
                           const double *Aci, double *PTA, size_t ldap, double *WTT,
                           hipStream_t s)
{
    hipLaunchKernelGGL(k_coarse_cols, dim3((n + 255) / 256, nc), dim3(256), 0, s, n, agg, G, H,
                       ld, alpha, gmask, constrained, PTA, ldap);
    hipLaunchKernelGGL(k_coarse_w, dim3((n + 255) / 256, (nc + CW_TILE - 1) / CW_TILE), dim3(256),
                       0, s, n, nc, (const double *)PTA, ldap, Aci, 1, WTT, (float *)nullptr);
}

void launch_coarse_pre(int is_float, int n, int agg, int nc, const void *X, size_t ldx,
                       const void *WT, size_t ldw, double *C, void *Tout, int k, hipStream_t s)
{
    const dim3 cg(nc, k), rg((n + 63) / 64, k);
    if (is_float) {
        hipLaunchKernelGGL(k_coarse_restrict<float>, cg, dim3(256), 0, s, n, agg,
                           (const float *)X, ldx, C);
        hipLaunchKernelGGL(k_coarse_residual<float>, rg, dim3(256), 0, s, n, nc, (const float *)X,
                           ldx, (const float *)WT, ldw, (const double *)C, (float *)Tout);
    } else {
        hipLaunchKernelGGL(k_coarse_restrict<double>, cg, dim3(256), 0, s, n, agg,
                           (const double *)X, ldx, C);
        hipLaunchKernelGGL(k_coarse_residual<double>, rg, dim3(256), 0, s, n, nc,
                           (const double *)X, ldx, (const double *)WT, ldw, (const double *)C,
                           (double *)Tout);
    }
}

void launch_coarse_restrict(int is_float, int n, int agg, int nc, const void *X, size_t ldx,
                            double *C, int k, hipStream_t s)
{
    const dim3 cg(nc, k);
    if (is_float)
        hipLaunchKernelGGL(k_coarse_restrict<float>, cg, dim3(256), 0, s, n, agg,
                           (const float *)X, ldx, C);
    else
        hipLaunchKernelGGL(k_coarse_restrict<double>, cg, dim3(256), 0, s, n, agg,
                           (const double *)X, ldx, C);
}

bool bj_coarse_fits(int bs, int agg, int nc) { return bs >= 4 && bs <= 64 && agg >= bs && nc <= CS_MAX; }

void launch_bj_coarse(int is_float, int n, int bs, const void *binv, const void *x, void *y,
                      const void *WT, size_t ldw, const double *C, const void *Aci, int nc,
                      int agg, hipStream_t s, const double *Cpart)
{
    const int nb = (n + bs - 1) / bs;
    if (is_float)
        hipLaunchKernelGGL(k_bj_coarse<float>, dim3(nb), dim3(64 * BJC_W), 0, s, n, bs,
                           (const float *)binv, (const float *)x, (float *)y, (const float *)WT,
                           ldw, C, (const float *)Aci, nc, agg, Cpart);
    else
        hipLaunchKernelGGL(k_bj_coarse<double>, dim3(nb), dim3(64 * BJC_W), 0, s, n, bs,
                           (const double *)binv, (const double *)x, (double *)y,
                           (const double *)WT, ldw, C, (const double *)Aci, nc, agg, Cpart);
}

}  // namespace wb
