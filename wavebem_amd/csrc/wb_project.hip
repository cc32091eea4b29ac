// wb_project.hip -- L2 projections onto the continuous Q1 space of the dofs, on the device.
//
// Replaces the UMFPACK mass-matrix solves of
//   ComputationalDomain::compute_normals        (source/computational_domain.cc:1525-1620),
//   NumericalTowingTank::compute_normals_at_nodes (source/numerical_towing_tank.cc:3358-3433),
//   BEMProblem::compute_surface_gradients       (source/bem_problem.cc:1154-1290),
//   FreeSurface::compute_potential_gradients    (source/free_surface.cc:9271-9400):
//   sum_j M_ij x_j = sum_c sum_q phi_i(q) f(q) JxW(q),  M_ij = sum_c sum_q phi_i phi_j JxW,
// with f = the cell normal (normals) or the surface gradient of the Q1 interpolant of u
// (gradients; plus n psi for the complete potential gradient), one scalar mass matrix for the three components (the reference's vector
// mass matrix couples only equal components).
//
// The mass matrix stays in element form: ten values per cell (the symmetric 4x4 block),
// applied row by row through the node -> cell incidence in a fixed order, so the product
// needs no atomics and is bit-reproducible.  The six right-hand sides (normal x,y,z and
// gradient x,y,z) are solved together by Jacobi-preconditioned CG; vectors are [n][6].
// Every kernel is HBM/latency-bound and tiny next to the assembly (n * 6 doubles per
// vector); the point is to keep the geometry on the device and the host out of the loop.
#include "wb_internal.h"
#include "wb_device.h"

namespace wb {

namespace {

constexpr int NCOL = 6;
constexpr int PB = 256;  // threads per block of the row kernels

// packed index of the symmetric 4x4 element block, a <= b
__host__ __device__ constexpr int msym(int a, int b)
{
    return a <= b ? a * 4 - a * (a - 1) / 2 + (b - a) : b * 4 - b * (b - 1) / 2 + (a - b);
}

// one thread per cell: element mass block (10) and element right-hand sides (4 x 6)
// With psi (dphi/dn at the dofs), the gradient columns take the complete potential
// gradient n psi + surface gradient of u (FreeSurface::compute_potential_gradients,
// source/free_surface.cc:9325-9344).
__global__ void k_proj_cells(const double *__restrict__ X, const int *__restrict__ cell_int,
                             const int8_t *__restrict__ orient, const double *__restrict__ u,
                             const double *__restrict__ psi, int nc, int nq,
                             const double *__restrict__ qpts, const double *__restrict__ qw,
                             double *__restrict__ emass, double *__restrict__ erhs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double v[4][3], uc[4] = {0, 0, 0, 0}, pc[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        const int d = cell_int[4 * c + k];
        for (int e = 0; e < 3; ++e) v[k][e] = X[3 * d + e];
        if (u) uc[k] = u[d];
        if (psi) pc[k] = psi[d];
    }
    const double so = orient[c] < 0 ? -1.0 : 1.0;
    double m[10], r[4][NCOL];
    for (int k = 0; k < 10; ++k) m[k] = 0.0;
    for (int k = 0; k < 4; ++k)
        for (int e = 0; e < NCOL; ++e) r[k][e] = 0.0;
    for (int q = 0; q < nq; ++q) {
        double sh[4], dx[4], de[4];
        q1_shape(qpts[2 * q], qpts[2 * q + 1], sh, dx, de);
        double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
        double ux = 0.0, ue = 0.0, pq = 0.0;
        for (int k = 0; k < 4; ++k) {
            for (int e = 0; e < 3; ++e) {
                a[e] += dx[k] * v[k][e];
                b[e] += de[k] * v[k][e];
            }
            ux += dx[k] * uc[k];
            ue += de[k] * uc[k];
            pq += sh[k] * pc[k];
        }
        const double g00 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        const double g01 = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
        const double g11 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        const double det = g00 * g11 - g01 * g01;
        const double jxw = sqrt(det) * qw[q];
        const double nn[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                              a[0] * b[1] - a[1] * b[0]};
        const double ns = so / sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
        // surface gradient of the interpolant: J (J^T J)^-1 (du/dxi, du/deta)
        const double ca = (g11 * ux - g01 * ue) / det, cb = (g00 * ue - g01 * ux) / det;
        double f[NCOL];
        for (int e = 0; e < 3; ++e) {
            f[e] = nn[e] * ns * jxw;
            f[3 + e] = (ca * a[e] + cb * b[e] + nn[e] * ns * pq) * jxw;
        }
        for (int i = 0; i < 4; ++i) {
            for (int j = i; j < 4; ++j) m[msym(i, j)] += sh[i] * sh[j] * jxw;
            for (int e = 0; e < NCOL; ++e) r[i][e] += sh[i] * f[e];
        }
    }
    for (int k = 0; k < 10; ++k) emass[10 * (size_t)c + k] = m[k];
    for (int i = 0; i < 4; ++i)
        for (int e = 0; e < NCOL; ++e) erhs[(size_t)(4 * c + i) * NCOL + e] = r[i][e];
}

// right-hand side and inverse diagonal per row, gathered over the incident cells
__global__ void k_proj_rows(const int *__restrict__ inc_ptr, const int *__restrict__ inc_item,
                            const double *__restrict__ emass, const double *__restrict__ erhs,
                            int n, double *__restrict__ B, double *__restrict__ dinv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double b[NCOL] = {0, 0, 0, 0, 0, 0}, d = 0.0;
    for (int p = inc_ptr[i]; p < inc_ptr[i + 1]; ++p) {
        const int it = inc_item[p], c = it >> 2, a = it & 3;
        d += emass[10 * (size_t)c + msym(a, a)];
        for (int e = 0; e < NCOL; ++e) b[e] += erhs[(size_t)it * NCOL + e];
    }
    for (int e = 0; e < NCOL; ++e) B[(size_t)i * NCOL + e] = b[e];
    dinv[i] = d > 0.0 ? 1.0 / d : 0.0;
}

// sum of v over the block (256 threads), fixed order; result valid in thread 0
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double (*lds)[K])
{
    for (int k = 0; k < K; ++k)
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < K; ++k) lds[wv][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < K; ++k) v[k] = ((lds[0][k] + lds[1][k]) + lds[2][k]) + lds[3][k];
}

// totals of the block partials part[nb][K] (fixed order), broadcast through LDS
template <int K>
__device__ __forceinline__ void sum_parts(const double *__restrict__ part, int nb, double *out)
{
    if (threadIdx.x < K) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += part[(size_t)b * K + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// Q = M P, partial P.Q
__global__ void __launch_bounds__(PB) k_proj_apply(
    const int *__restrict__ inc_ptr, const int *__restrict__ inc_item,
    const int *__restrict__ cell_int, const double *__restrict__ emass,
    const double *__restrict__ P, int n, double *__restrict__ Q, double *__restrict__ part)
{
    __shared__ double lds[PB / 64][NCOL];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[NCOL] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        double q[NCOL] = {0, 0, 0, 0, 0, 0};
        for (int p = inc_ptr[i]; p < inc_ptr[i + 1]; ++p) {
            const int it = inc_item[p], c = it >> 2, a = it & 3;
            const double *m = emass + 10 * (size_t)c;
            for (int b = 0; b < 4; ++b) {
                const double w = m[msym(a, b)];
                const double *pj = P + (size_t)cell_int[4 * c + b] * NCOL;
                for (int e = 0; e < NCOL; ++e) q[e] += w * pj[e];
            }
        }
        for (int e = 0; e < NCOL; ++e) {
            Q[(size_t)i * NCOL + e] = q[e];
            acc[e] = P[(size_t)i * NCOL + e] * q[e];
        }
    }
    block_sum<NCOL>(acc, lds);
    if (threadIdx.x == 0)
        for (int e = 0; e < NCOL; ++e) part[(size_t)blockIdx.x * NCOL + e] = acc[e];
}

// state: rho[2][6] (by iteration parity), bb[6], rr[6]
// alpha = rho / (p.q); x += alpha p; r -= alpha q; z = D^-1 r; partials r.z, r.r
__global__ void __launch_bounds__(PB) k_proj_update(
    const double *__restrict__ part, int nb, const double *__restrict__ state, int iter,
    const double *__restrict__ P, const double *__restrict__ Q, const double *__restrict__ dinv,
    int n, double *__restrict__ Xs, double *__restrict__ R, double *__restrict__ Z,
    double *__restrict__ part2)
{
    __shared__ double tot[NCOL];
    __shared__ double lds[PB / 64][2 * NCOL];
    sum_parts<NCOL>(part, nb, tot);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[2 * NCOL];
    for (int k = 0; k < 2 * NCOL; ++k) acc[k] = 0.0;
    if (i < n) {
        const double di = dinv[i];
        for (int e = 0; e < NCOL; ++e) {
            const double rho = state[(iter & 1) * NCOL + e], pq = tot[e];
            const double alpha = pq > 0.0 ? rho / pq : 0.0;
            const size_t o = (size_t)i * NCOL + e;
            Xs[o] += alpha * P[o];
            const double r = R[o] - alpha * Q[o];
            R[o] = r;
            const double z = di * r;
            Z[o] = z;
            acc[e] = r * z;
            acc[NCOL + e] = r * r;
        }
    }
    block_sum<2 * NCOL>(acc, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 2 * NCOL; ++k) part2[(size_t)blockIdx.x * 2 * NCOL + k] = acc[k];
}

// beta = rho_new / rho; p = z + beta p (first call: p = z); records rho_new and r.r
__global__ void __launch_bounds__(PB) k_proj_dir(
    const double *__restrict__ part2, int nb, double *__restrict__ state, int iter,
    const double *__restrict__ Z, int n, double *__restrict__ P)
{
    __shared__ double tot[2 * NCOL];
    sum_parts<2 * NCOL>(part2, nb, tot);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool first = iter < 0;
    if (i < n)
        for (int e = 0; e < NCOL; ++e) {
            const size_t o = (size_t)i * NCOL + e;
            if (first) {
                P[o] = Z[o];
            } else {
                const double rho = state[(iter & 1) * NCOL + e];
                const double beta = rho > 0.0 ? tot[e] / rho : 0.0;
                P[o] = Z[o] + beta * P[o];
            }
        }
    if (blockIdx.x == 0 && threadIdx.x < NCOL) {
        const int e = threadIdx.x;
        state[((iter + 1) & 1) * NCOL + e] = tot[e];
        if (first) state[2 * NCOL + e] = tot[NCOL + e];
        state[3 * NCOL + e] = tot[NCOL + e];
    }
}

// x = 0, r = b, z = D^-1 b, partials r.z, r.r
__global__ void __launch_bounds__(PB) k_proj_init(const double *__restrict__ B,
                                                  const double *__restrict__ dinv, int n,
                                                  double *__restrict__ Xs, double *__restrict__ R,
                                                  double *__restrict__ Z,
                                                  double *__restrict__ part2)
{
    __shared__ double lds[PB / 64][2 * NCOL];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[2 * NCOL];
    for (int k = 0; k < 2 * NCOL; ++k) acc[k] = 0.0;
    if (i < n) {
        const double di = dinv[i];
        for (int e = 0; e < NCOL; ++e) {
            const size_t o = (size_t)i * NCOL + e;
            const double r = B[o];
            Xs[o] = 0.0;
            R[o] = r;
            Z[o] = di * r;
            acc[e] = r * di * r;
            acc[NCOL + e] = r * r;
        }
    }
    block_sum<2 * NCOL>(acc, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 2 * NCOL; ++k) part2[(size_t)blockIdx.x * 2 * NCOL + k] = acc[k];
}

}  // namespace

int proj_blocks(int n) { return (n + PB - 1) / PB; }

void launch_proj_setup(const double *X, const int *cell_int, const int8_t *orient,
                       const double *u, const double *psi, int nc, int nq, const double *qpts,
                       const double *qw, const int *inc_ptr, const int *inc_item, int n,
                       double *emass, double *erhs, double *B, double *dinv, hipStream_t s)
{
    hipLaunchKernelGGL(k_proj_cells, dim3((nc + 127) / 128), dim3(128), 0, s, X, cell_int,
                       orient, u, psi, nc, nq, qpts, qw, emass, erhs);
    hipLaunchKernelGGL(k_proj_rows, dim3(proj_blocks(n)), dim3(PB), 0, s, inc_ptr, inc_item,
                       emass, erhs, n, B, dinv);
}

void launch_proj_cg_start(const double *B, const double *dinv, int n, double *Xs, double *R,
                          double *Z, double *P, double *part2, double *state, hipStream_t s)
{
    const int nb = proj_blocks(n);
    hipLaunchKernelGGL(k_proj_init, dim3(nb), dim3(PB), 0, s, B, dinv, n, Xs, R, Z, part2);
    hipLaunchKernelGGL(k_proj_dir, dim3(nb), dim3(PB), 0, s, part2, nb, state, -1, Z, n, P);
}

void launch_proj_cg_iter(const int *inc_ptr, const int *inc_item, const int *cell_int,
                         const double *emass, const double *dinv, int n, int iter, double *Xs,
                         double *R, double *Z, double *P, double *Q, double *part,
                         double *part2, double *state, hipStream_t s)
{
    const int nb = proj_blocks(n);
    hipLaunchKernelGGL(k_proj_apply, dim3(nb), dim3(PB), 0, s, inc_ptr, inc_item, cell_int,
                       emass, P, n, Q, part);
    hipLaunchKernelGGL(k_proj_update, dim3(nb), dim3(PB), 0, s, part, nb, state, iter, P, Q,
                       dinv, n, Xs, R, Z, part2);
    hipLaunchKernelGGL(k_proj_dir, dim3(nb), dim3(PB), 0, s, part2, nb, state, iter, Z, n, P);
}

}  // namespace wb
