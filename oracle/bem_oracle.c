/*
 * bem_oracle.c -- CPU oracle for WaveBEM's collocation-BEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see bem_oracle.h): used by tests/, the smoke check and
 * the bench's cpu_baseline leg as the checker.  The product never links it.
 *
 * Compiled with -ffp-contract=off so that every expression rounds as written, in the
 * evaluation order of the reference expressions cited below.
 */
#include "bem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------------------------
 * deal.II QGauss<1>(n) on [0,1]: Newton iteration on the Legendre recurrence in long
 * double, nodes symmetric about 1/2, ascending (deal.II base/quadrature_lib.cc,
 * QGauss<1>::QGauss; called through QuadratureSelector("gauss", 4),
 * source/computational_domain.cc:125-128).
 * ---------------------------------------------------------------------------------- */
int orc_qgauss_1d(int n, double *x, double *w)
{
    if (n <= 0) return -1;
    const int m = (n + 1) / 2;
    volatile long double one = 1.0L;
    const long double leps = (long double)__LDBL_EPSILON__;
    const long double deps = (long double)__DBL_EPSILON__;
    const long double tol = (one + leps != one) ? fmaxl(deps / 100, leps * 5) : deps * 5;
    for (int i = 1; i <= m; ++i) {
        long double z = cosl((long double)ORC_PI * (i - .25L) / (n + .5L));
        long double pp, p1, p2, p3;
        do {
            p1 = 1.L;
            p2 = 0.L;
            for (int j = 0; j < n; ++j) {
                p3 = p2;
                p2 = p1;
                p1 = ((2.L * j + 1.L) * z * p2 - j * p3) / (j + 1);
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            z = z - p1 / pp;
        } while (fabsl(p1 / pp) > tol);
        const double xx = (double)(.5L * z);
        const double ww = (double)(1.L / ((1.L - z * z) * pp * pp));
        x[i - 1] = .5 - xx;
        x[n - i] = .5 + xx;
        w[i - 1] = ww;
        w[n - i] = ww;
    }
    return 0;
}

/* QGauss<2>(n): tensor product, first coordinate fastest (deal.II Quadrature<dim>
 * tensor-product constructor). pts[2*q+{0,1}] */
int orc_qgauss_2d(int n, double *pts, double *w)
{
    double x[64], wx[64];
    if (n <= 0 || n > 64) return -1;
    orc_qgauss_1d(n, x, wx);
    int q = 0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i, ++q) {
            pts[2 * q + 0] = x[i];
            pts[2 * q + 1] = x[j];
            w[q] = wx[i] * wx[j];
        }
    return 0;
}

/* QGaussOneOverR<2>(n, vertex, factor_out): 2 n^2 points.  Polar-type map of the
 * QGauss<2>(n) rule onto the triangle below the diagonal from vertex 0,
 * (u,v) -> (u, u tan(pi/4 v)) with weight w pi/4 / cos(pi/4 v), times |p| when the 1/R
 * singularity is factored out; the mirror triangle repeats the weights with swapped
 * coordinates; other vertices rotate the rule about (1/2,1/2) by pi/2, -pi/2, pi
 * (deal.II base/quadrature_lib.cc, QGaussOneOverR<2>).  The reference builds it with
 * the vertex support point (source/bem_problem.cc:116-120); for a vertex singularity
 * deal.II's point constructor keeps exactly one unit box, which reduces to this rule. */
int orc_qgauss_one_over_r(int n, int vertex, int factor_out, double *pts, double *w)
{
    if (vertex < 0 || vertex > 3 || n <= 0 || n > 64) return -1;
    double *gp = (double *)malloc(sizeof(double) * 2 * n * n);
    double *gw = (double *)malloc(sizeof(double) * n * n);
    orc_qgauss_2d(n, gp, gw);
    const int ng = n * n;
    const double pi4 = ORC_PI / 4;
    for (int q = 0; q < ng; ++q) {
        const double u = gp[2 * q], v = gp[2 * q + 1];
        double p0 = u;
        double p1 = u * tan(pi4 * v);
        double ww = gw[q] * pi4 / cos(pi4 * v);
        if (factor_out) ww *= sqrt(p0 * p0 + p1 * p1);
        pts[2 * q + 0] = p0;
        pts[2 * q + 1] = p1;
        w[q] = ww;
        pts[2 * (ng + q) + 0] = p1;
        pts[2 * (ng + q) + 1] = p0;
        w[ng + q] = ww;
    }
    double theta = 0;
    switch (vertex) {
    case 1: theta = ORC_PI / 2; break;
    case 2: theta = -ORC_PI / 2; break;
    case 3: theta = ORC_PI; break;
    default: theta = 0;
    }
    const double R00 = cos(theta), R01 = -sin(theta);
    const double R10 = sin(theta), R11 = cos(theta);
    if (vertex != 0)
        for (int q = 0; q < 2 * ng; ++q) {
            const double x = pts[2 * q] - .5, y = pts[2 * q + 1] - .5;
            pts[2 * q + 0] = R00 * x + R01 * y + .5;
            pts[2 * q + 1] = R10 * x + R11 * y + .5;
        }
    free(gp);
    free(gw);
    return 0;
}

/* Q1 shape functions in deal.II lexicographic order v0=(0,0) v1=(1,0) v2=(0,1) v3=(1,1) */
static void q1_shape(double xi, double eta, double *phi, double *dxi, double *deta)
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

/* FEValues<2,3>(MappingQ1Eulerian, FE_Q(1), quadrature) on one cell: quadrature points
 * x = sum_k phi_k v_k; Jacobian columns a = sum dphi_k/dxi v_k, b = sum dphi_k/deta v_k;
 * JxW = sqrt(det(J^T J)) * w; normal = (a x b)/|a x b|, negated when the cell direction
 * flag is false (deal.II MappingQGeneric::fill_fe_values for codimension one).  The
 * reference requests them at source/bem_problem.cc:133-137 and :495-501. */
void orc_fe_values(const double *v, int orient, int nq, const double *pts, const double *w,
                   double *qp, double *normals, double *jxw, double *shape)
{
    for (int q = 0; q < nq; ++q) {
        double phi[4], dxi[4], deta[4];
        q1_shape(pts[2 * q], pts[2 * q + 1], phi, dxi, deta);
        double x[3] = {0, 0, 0}, a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
        for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) {
                x[d] += phi[k] * v[3 * k + d];
                a[d] += dxi[k] * v[3 * k + d];
                b[d] += deta[k] * v[3 * k + d];
            }
        const double g00 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        const double g01 = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
        const double g11 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        const double det = g00 * g11 - g01 * g01;
        jxw[q] = sqrt(det) * w[q];
        double nn[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                        a[0] * b[1] - a[1] * b[0]};
        const double nl = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
        for (int d = 0; d < 3; ++d) {
            nn[d] /= nl;
            if (orient < 0) nn[d] *= -1.;
            qp[3 * q + d] = x[d];
            normals[3 * q + d] = nn[d];
        }
        for (int k = 0; k < 4; ++k) shape[4 * q + k] = phi[k];
    }
}

/* LaplaceKernel::kernels<3> (include/laplace_kernel.h:45-64):
 * r=|R|, d = 1/(r*4*pi), D = R/(-4*pi*r^2*r). */
static inline void laplace_kernels(const double *R, double *D, double *d)
{
    const double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    const double r2 = r * r;
    *d = (1. / (r * 4 * ORC_PI));
    const double den = -4 * ORC_PI * r2 * r;
    D[0] = R[0] / den;
    D[1] = R[1] / den;
    D[2] = R[2] / den;
}

static inline int in_set(const int *dn_ptr, const int *dn_idx, int i, int j)
{
    for (int p = dn_ptr[i]; p < dn_ptr[i + 1]; ++p)
        if (dn_idx[p] == j) return 1;
    return 0;
}

/* BEMProblem::assemble_system (source/bem_problem.cc:106-590), Direct branch.
 * The reference loops cells (outer) then rows i (inner).  Here rows are the outer loop
 * (so rows can run in parallel) and cells the inner one: every matrix entry still
 * receives its per-cell contributions in cell order, and each per-cell local value is
 * summed over q in the same order, so the result is identical to the reference order.
 * Per-cell FEValues (regular rule) and the four singular FEValues (one per vertex) are
 * evaluated once per cell; they depend only on the cell. */
int orc_assemble(int n, int nc, const int *cell_dofs, const signed char *orient,
                 const double *support, const int *dn_ptr, const int *dn_idx, int q_order,
                 int sing_order, double *H, double *G, int n_threads)
{
    return orc_assemble_rows(n, nc, cell_dofs, orient, support, dn_ptr, dn_idx, q_order,
                             sing_order, 0, n, H, G, n_threads);
}

/* Rows [row0, row1) of assemble_system into H, G of shape (row1-row0) x n. */
int orc_assemble_rows(int n, int nc, const int *cell_dofs, const signed char *orient,
                      const double *support, const int *dn_ptr, const int *dn_idx, int q_order,
                      int sing_order, int row0, int row1, double *H, double *G, int n_threads)
{
    if (n <= 0 || nc <= 0 || q_order <= 0 || q_order > 16 || sing_order <= 0 ||
        sing_order > 16 || row0 < 0 || row1 > n || row1 < row0)
        return -1;
    const int nq = q_order * q_order;
    const int ns = 2 * sing_order * sing_order;
    double *qpts = (double *)malloc(sizeof(double) * 2 * nq);
    double *qw = (double *)malloc(sizeof(double) * nq);
    orc_qgauss_2d(q_order, qpts, qw);
    double *spts = (double *)malloc(sizeof(double) * 4 * 2 * ns);
    double *sw = (double *)malloc(sizeof(double) * 4 * ns);
    for (int k = 0; k < 4; ++k)
        orc_qgauss_one_over_r(sing_order, k, 1, spts + 2 * ns * k, sw + ns * k);

    /* per cell: regular rule data, then 4 singular rules */
    const size_t per_reg = (size_t)nq * 11;     /* qp3 n3 jxw1 shape4 */
    const size_t per_sing = (size_t)4 * ns * 11;
    double *creg = (double *)malloc(sizeof(double) * per_reg * nc);
    double *csing = (double *)malloc(sizeof(double) * per_sing * nc);
    if (!qpts || !qw || !spts || !sw || !creg || !csing) return -2;
    for (int c = 0; c < nc; ++c) {
        double v[12];
        for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) v[3 * k + d] = support[3 * cell_dofs[4 * c + k] + d];
        double *r = creg + per_reg * c;
        orc_fe_values(v, orient[c], nq, qpts, qw, r, r + 3 * nq, r + 6 * nq, r + 7 * nq);
        for (int k = 0; k < 4; ++k) {
            double *s = csing + per_sing * c + (size_t)11 * ns * k;
            orc_fe_values(v, orient[c], ns, spts + 2 * ns * k, sw + ns * k, s, s + 3 * ns,
                          s + 6 * ns, s + 7 * ns);
        }
    }
    memset(H, 0, sizeof(double) * (size_t)(row1 - row0) * n);
    memset(G, 0, sizeof(double) * (size_t)(row1 - row0) * n);
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int i = row0; i < row1; ++i) {
        const double *xi = support + 3 * i;
        double *Hi = H + (size_t)(i - row0) * n;
        double *Gi = G + (size_t)(i - row0) * n;
        for (int c = 0; c < nc; ++c) {
            const int *dofs = cell_dofs + 4 * c;
            double lH[4] = {0, 0, 0, 0}, lG[4] = {0, 0, 0, 0};
            int sing = -1;
            for (int j = 0; j < 4; ++j)
                if (in_set(dn_ptr, dn_idx, i, dofs[j])) {
                    sing = j;
                    break;
                }
            const double *base;
            int nqq;
            if (sing < 0) {
                base = creg + per_reg * c;
                nqq = nq;
            } else {
                base = csing + per_sing * c + (size_t)11 * ns * sing;
                nqq = ns;
            }
            const double *qp = base, *nr = base + 3 * nqq, *jw = base + 6 * nqq,
                         *sh = base + 7 * nqq;
            for (int q = 0; q < nqq; ++q) {
                const double R[3] = {qp[3 * q] + -1.0 * xi[0], qp[3 * q + 1] + -1.0 * xi[1],
                                     qp[3 * q + 2] + -1.0 * xi[2]};
                double D[3], s;
                laplace_kernels(R, D, &s);
                const double Dn = D[0] * nr[3 * q] + D[1] * nr[3 * q + 1] + D[2] * nr[3 * q + 2];
                for (int j = 0; j < 4; ++j) {
                    lH[j] += ((Dn * sh[4 * q + j]) * jw[q]);
                    lG[j] += ((s * sh[4 * q + j]) * jw[q]);
                }
            }
            for (int j = 0; j < 4; ++j) {
                Hi[dofs[j]] += lH[j];
                Gi[dofs[j]] += lG[j];
            }
        }
    }
    free(qpts);
    free(qw);
    free(spts);
    free(sw);
    free(creg);
    free(csing);
    return 0;
}

/* deal.II FullMatrix::vmult / vmult_add: s = (adding ? dst_i : 0); s += M_ij src_j */
static void fm_vmult(int n, const double *M, const double *src, double *dst, int adding)
{
    for (int i = 0; i < n; ++i) {
        const double *row = M + (size_t)i * n;
        double s = adding ? dst[i] : 0.;
        for (int j = 0; j < n; ++j) s += src[j] * row[j];
        dst[i] = s;
    }
}

/* Seconds per FullMatrix::vmult of an n x n matrix, rows split over n_threads (the
 * per-row arithmetic is that of fm_vmult); used only to time the CPU baseline. */
double orc_gemv_timed(int n, const double *M, const double *x, double *y, int reps,
                      int n_threads)
{
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
    double t0 = omp_get_wtime();
#pragma omp parallel
    for (int r = 0; r < reps; ++r) {
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            const double *row = M + (size_t)i * n;
            double s = 0.;
            for (int j = 0; j < n; ++j) s += x[j] * row[j];
            y[i] = s;
        }
    }
    return (omp_get_wtime() - t0) / (reps > 0 ? reps : 1);
#else
    (void)n_threads;
    for (int r = 0; r < reps; ++r) fm_vmult(n, M, x, y, 0);
    return -1.0;
#endif
}

/* compute_alpha (source/bem_problem.cc:594-618): alpha = H * (-1) */
void orc_compute_alpha(int n, const double *H, double *alpha)
{
    double *ones = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) ones[i] = -1.;
    fm_vmult(n, H, ones, alpha, 0);
    free(ones);
}

static double l2_norm(int n, const double *v)
{
    double s = 0;
    for (int i = 0; i < n; ++i) s += v[i] * v[i];
    return sqrt(s);
}

/* BEMProblem::vmult (source/bem_problem.cc:621-670), Direct branch */
void orc_vmult(int n, const double *G, const double *H, const double *alpha,
               const double *surface, const double *other, const double *src, double *dst)
{
    double *phi = (double *)malloc(sizeof(double) * n);
    double *dphi = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
        phi[i] = src[i] * other[i];
        dphi[i] = src[i] * surface[i];
    }
    fm_vmult(n, G, dphi, dst, 0);
    for (int i = 0; i < n; ++i) dst[i] *= -1;
    fm_vmult(n, H, phi, dst, 1);
    for (int i = 0; i < n; ++i) dst[i] += phi[i] * alpha[i];
    double linf = 0;
    for (int i = 0; i < n; ++i) linf = fmax(linf, fabs(surface[i]));
    if (linf < 1e-10) {
        const double nrm = l2_norm(n, dst);
        for (int i = 0; i < n; ++i) dst[i] += -nrm;
    }
    free(phi);
    free(dphi);
}

/* BEMProblem::compute_rhs (source/bem_problem.cc:673-717), Direct branch */
void orc_compute_rhs(int n, const double *G, const double *H, const double *alpha,
                     const double *surface, const double *other, const double *src, double *dst)
{
    double *phi = (double *)malloc(sizeof(double) * n);
    double *dphi = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
        phi[i] = src[i] * surface[i];
        dphi[i] = src[i] * other[i];
    }
    fm_vmult(n, H, phi, dst, 0);
    for (int i = 0; i < n; ++i) dst[i] += phi[i] * alpha[i];
    for (int i = 0; i < n; ++i) dst[i] *= -1;
    fm_vmult(n, G, dphi, dst, 1);
    free(phi);
    free(dphi);
}

/* ---------------- ConstraintMatrix (deal.II) with the rules of compute_constraints ---- */
typedef struct {
    int row;
    int n, cap;
    int *col;
    double *w;
    double inhom;
} cline;

typedef struct {
    int n;
    int *line_of; /* dof -> line index or -1 */
    cline *lines;
    int nl, capl;
} cmatrix;

static void cm_init(cmatrix *c, int n)
{
    c->n = n;
    c->line_of = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) c->line_of[i] = -1;
    c->nl = 0;
    c->capl = 16;
    c->lines = (cline *)malloc(sizeof(cline) * c->capl);
}

static void cm_free(cmatrix *c)
{
    for (int l = 0; l < c->nl; ++l) {
        free(c->lines[l].col);
        free(c->lines[l].w);
    }
    free(c->lines);
    free(c->line_of);
}

/* ConstraintMatrix::add_line: no-op when the line exists */
static void cm_add_line(cmatrix *c, int row)
{
    if (c->line_of[row] >= 0) return;
    if (c->nl == c->capl) {
        c->capl *= 2;
        c->lines = (cline *)realloc(c->lines, sizeof(cline) * c->capl);
    }
    cline *L = &c->lines[c->nl];
    L->row = row;
    L->n = 0;
    L->cap = 4;
    L->col = (int *)malloc(sizeof(int) * 4);
    L->w = (double *)malloc(sizeof(double) * 4);
    L->inhom = 0;
    c->line_of[row] = c->nl++;
}

/* ConstraintMatrix::add_entry: an existing (line, column) entry is kept as is */
static void cm_add_entry(cmatrix *c, int row, int col, double w)
{
    cline *L = &c->lines[c->line_of[row]];
    for (int k = 0; k < L->n; ++k)
        if (L->col[k] == col) return;
    if (L->n == L->cap) {
        L->cap *= 2;
        L->col = (int *)realloc(L->col, sizeof(int) * L->cap);
        L->w = (double *)realloc(L->w, sizeof(double) * L->cap);
    }
    L->col[L->n] = col;
    L->w[L->n] = w;
    L->n++;
}

static void cm_set_inhom(cmatrix *c, int row, double v) { c->lines[c->line_of[row]].inhom = v; }

/* ConstraintMatrix::close: substitute entries that point at constrained dofs (adding
 * their inhomogeneity), merge duplicate columns, sort entries and lines. */
static int cm_close(cmatrix *c)
{
    for (int pass = 0; pass < 64; ++pass) {
        int changed = 0;
        for (int l = 0; l < c->nl; ++l) {
            cline *L = &c->lines[l];
            for (int k = 0; k < L->n; ++k) {
                const int j = L->col[k];
                const int lj = c->line_of[j];
                if (lj < 0) continue;
                if (j == L->row) return -3; /* self reference */
                const double wk = L->w[k];
                cline *M = &c->lines[lj];
                L->inhom += wk * M->inhom;
                /* remove entry k */
                L->col[k] = L->col[L->n - 1];
                L->w[k] = L->w[L->n - 1];
                L->n--;
                for (int m = 0; m < M->n; ++m) {
                    int found = 0;
                    for (int t = 0; t < L->n; ++t)
                        if (L->col[t] == M->col[m]) {
                            L->w[t] += wk * M->w[m];
                            found = 1;
                            break;
                        }
                    if (!found) {
                        if (L->n == L->cap) {
                            L->cap *= 2;
                            L->col = (int *)realloc(L->col, sizeof(int) * L->cap);
                            L->w = (double *)realloc(L->w, sizeof(double) * L->cap);
                        }
                        L->col[L->n] = M->col[m];
                        L->w[L->n] = wk * M->w[m];
                        L->n++;
                    }
                }
                changed = 1;
                k = -1;
            }
        }
        if (!changed) break;
        if (pass == 63) return -4; /* cycle */
    }
    /* sort entries by column (insertion sort, lines are short) */
    for (int l = 0; l < c->nl; ++l) {
        cline *L = &c->lines[l];
        for (int a = 1; a < L->n; ++a) {
            int cc = L->col[a];
            double ww = L->w[a];
            int b = a - 1;
            while (b >= 0 && L->col[b] > cc) {
                L->col[b + 1] = L->col[b];
                L->w[b + 1] = L->w[b];
                --b;
            }
            L->col[b + 1] = cc;
            L->w[b + 1] = ww;
        }
    }
    return 0;
}

static int cmp_line_row(const void *a, const void *b)
{
    return ((const cline *)a)->row - ((const cline *)b)->row;
}

/* BEMProblem::compute_constraints (source/bem_problem.cc:990-1105).  Hanging-node
 * constraints (:1005) are not produced: meshes handed in here are conforming. */
int orc_compute_constraints(int n, const int *dn_ptr, const int *dn_idx, const double *surface,
                            const double *normals, const double *grads, const double *tmp_rhs,
                            int cap_lines, int cap_nnz, int *line_rows, int *line_ptr,
                            int *ent_col, double *ent_w, double *inhom)
{
    cmatrix c;
    cm_init(&c, n);
    int err = 0;
    for (int i = 0; i < n && !err; ++i) {
        const int p0 = dn_ptr[i], p1 = dn_ptr[i + 1];
        int first = dn_idx[p0];
        for (int p = p0; p < p1; ++p)
            if (surface[dn_idx[p]] == 1) {
                first = dn_idx[p];
                break;
            }
        if (i != first) continue;
        if (surface[i] == 1) {
            for (int p = p0; p < p1; ++p) {
                const int k = dn_idx[p];
                if (k == i) continue;
                if (surface[k] == 1) {
                    if (!normals) {
                        err = -5;
                        break;
                    }
                    const double *ni = normals + 3 * i, *nk = normals + 3 * k;
                    const double dx = nk[0] - ni[0], dy = nk[1] - ni[1], dz = nk[2] - ni[2];
                    if (sqrt(dx * dx + dy * dy + dz * dz) < 1e-4) {
                        cm_add_line(&c, k);
                        cm_add_entry(&c, k, i, 1);
                    } else {
                        if (!grads) {
                            err = -6;
                            break;
                        }
                        const double *gi = grads + 3 * i, *gk = grads + 3 * k;
                        const double nin = ni[0] * nk[0] + ni[1] * nk[1] + ni[2] * nk[2];
                        const double gkni = gk[0] * ni[0] + gk[1] * ni[1] + gk[2] * ni[2];
                        const double gink = gi[0] * nk[0] + gi[1] * nk[1] + gi[2] * nk[2];
                        const double this_ng = (1.0 / (1.0 - pow(nin, 2))) * (gkni + gink * nin);
                        const double other_ng = (1.0 / (1.0 - pow(nin, 2))) * (gink + gkni * nin);
                        cm_add_line(&c, k);
                        cm_add_line(&c, i);
                        cm_set_inhom(&c, i, this_ng);
                        cm_add_line(&c, k);
                        cm_set_inhom(&c, k, other_ng);
                    }
                } else {
                    cm_add_line(&c, k);
                    cm_set_inhom(&c, k, tmp_rhs[i]);
                }
            }
        } else {
            for (int p = p0; p < p1; ++p) {
                const int k = dn_idx[p];
                if (k == i) continue;
                cm_add_line(&c, k);
                cm_add_entry(&c, k, i, 1);
            }
        }
    }
    if (!err) err = cm_close(&c);
    if (err) {
        cm_free(&c);
        return err;
    }
    qsort(c.lines, c.nl, sizeof(cline), cmp_line_row);
    int nnz = 0;
    for (int l = 0; l < c.nl; ++l) nnz += c.lines[l].n;
    if (c.nl > cap_lines || nnz > cap_nnz) {
        cm_free(&c);
        return -7;
    }
    int pos = 0;
    line_ptr[0] = 0;
    for (int l = 0; l < c.nl; ++l) {
        line_rows[l] = c.lines[l].row;
        inhom[l] = c.lines[l].inhom;
        for (int k = 0; k < c.lines[l].n; ++k) {
            ent_col[pos] = c.lines[l].col[k];
            ent_w[pos] = c.lines[l].w[k];
            ++pos;
        }
        line_ptr[l + 1] = pos;
    }
    const int nl = c.nl;
    /* line_of was indexed by the pre-sort order; only free now */
    cm_free(&c);
    return nl;
}

/* ConstrainedOperator::vmult (include/constrained_matrix.h:73-86) over BEMProblem::vmult */
void orc_constrained_vmult(int n, const double *G, const double *H, const double *alpha,
                           const double *surface, const double *other, int n_lines,
                           const int *line_rows, const int *line_ptr, const int *ent_col,
                           const double *ent_w, const double *src, double *dst)
{
    orc_vmult(n, G, H, alpha, surface, other, src, dst);
    for (int l = 0; l < n_lines; ++l) {
        const int i = line_rows[l];
        dst[i] = src[i];
        for (int p = line_ptr[l]; p < line_ptr[l + 1]; ++p) dst[i] -= ent_w[p] * src[ent_col[p]];
    }
}

/* ------------- banded LU with partial pivoting (stands in for UMFPACK on the band) ---- */
typedef struct {
    int n, kl, ku, ld; /* storage: rows kl + (kl+ku) + 1, LAPACK gbtrf layout */
    double *ab;
    int *ipiv;
} band_lu;

#define AB(L, i, j) ((L)->ab[(size_t)(j) * (L)->ld + ((L)->kl + (L)->ku + (i) - (j))])

static int band_lu_factor(band_lu *L)
{
    const int n = L->n, kl = L->kl, ku = L->ku;
    const int kv = ku + kl; /* upper bandwidth after fill */
    for (int j = 0; j < n; ++j) {
        /* pivot search in column j, rows j..min(n-1, j+kl) */
        const int imax = j + kl < n - 1 ? j + kl : n - 1;
        int p = j;
        double best = fabs(AB(L, j, j));
        for (int i = j + 1; i <= imax; ++i)
            if (fabs(AB(L, i, j)) > best) {
                best = fabs(AB(L, i, j));
                p = i;
            }
        L->ipiv[j] = p;
        if (best == 0.0) return -1;
        const int jmax = j + kv < n - 1 ? j + kv : n - 1;
        if (p != j)
            for (int c = j; c <= jmax; ++c) {
                double t = AB(L, j, c);
                AB(L, j, c) = AB(L, p, c);
                AB(L, p, c) = t;
            }
        const double piv = AB(L, j, j);
        for (int i = j + 1; i <= imax; ++i) {
            const double m = AB(L, i, j) / piv;
            AB(L, i, j) = m;
            if (m != 0.0)
                for (int c = j + 1; c <= jmax; ++c) AB(L, i, c) -= m * AB(L, j, c);
        }
    }
    return 0;
}

static void band_lu_solve(const band_lu *L, double *x)
{
    const int n = L->n, kl = L->kl, ku = L->ku, kv = kl + ku;
    for (int j = 0; j < n; ++j) {
        const int p = L->ipiv[j];
        if (p != j) {
            double t = x[j];
            x[j] = x[p];
            x[p] = t;
        }
        const int imax = j + kl < n - 1 ? j + kl : n - 1;
        for (int i = j + 1; i <= imax; ++i) x[i] -= AB((band_lu *)L, i, j) * x[j];
    }
    for (int j = n - 1; j >= 0; --j) {
        x[j] /= AB((band_lu *)L, j, j);
        const int imin = j - kv > 0 ? j - kv : 0;
        for (int i = imin; i < j; ++i) x[i] -= AB((band_lu *)L, i, j) * x[j];
    }
}

/* assemble_preconditioner (source/bem_problem.cc:1107-1149): column i of the band holds
 * H(j,i) (+alpha_i on the diagonal) for a Neumann column, -G(j,i) for a Dirichlet column,
 * rows j in [i-band/2, i+band/2); constrained rows are identity rows. */
static int band_precond_build(band_lu *L, int n, int band, const double *G, const double *H,
                              const double *alpha, const double *surface,
                              const char *constrained)
{
    const int hb = band / 2;
    L->n = n;
    L->kl = hb - 1; /* row j - col i >= -(hb) ... */
    /* entries (j,i) with j in [i-hb, i+hb-1]: j - i in [-hb, hb-1] */
    L->kl = hb - 1;
    L->ku = hb;
    L->ld = 2 * L->kl + L->ku + 1;
    L->ab = (double *)calloc((size_t)L->ld * n, sizeof(double));
    L->ipiv = (int *)malloc(sizeof(int) * n);
    if (!L->ab || !L->ipiv) return -2;
    for (int i = 0; i < n; ++i) {
        if (constrained[i]) AB(L, i, i) = 1;
        const int j0 = i - hb > 0 ? i - hb : 0, j1 = i + hb < n ? i + hb : n;
        for (int j = j0; j < j1; ++j) {
            if (constrained[j]) continue;
            if (surface[i] == 0) {
                AB(L, j, i) = H[(size_t)j * n + i];
                if (i == j) AB(L, j, i) += alpha[i];
            } else
                AB(L, j, i) = -G[(size_t)j * n + i];
        }
    }
    return band_lu_factor(L);
}

typedef struct {
    int n;
    const double *G, *H, *alpha, *surface, *other;
    int n_lines;
    const int *line_rows, *line_ptr, *ent_col;
    const double *ent_w;
    int precond;
    band_lu lu;
} gm_op;

static void op_apply(const gm_op *o, const double *src, double *dst)
{
    orc_constrained_vmult(o->n, o->G, o->H, o->alpha, o->surface, o->other, o->n_lines,
                          o->line_rows, o->line_ptr, o->ent_col, o->ent_w, src, dst);
}

static void prec_apply(const gm_op *o, const double *src, double *dst)
{
    if (dst != src) memcpy(dst, src, sizeof(double) * o->n);
    if (o->precond == 1) band_lu_solve(&o->lu, dst);
}

static double dot(int n, const double *a, const double *b)
{
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* SolverControl::check: success if value <= tol, failure if step >= max or NaN */
static int sc_check(int step, double value, double tol, int max_steps)
{
    if (value <= tol) return 1;
    if (isnan(value) || step >= max_steps) return -1;
    return 0;
}

/* deal.II SolverGMRES<Vector<double>>::solve, left preconditioning, restart after
 * n_tmp-2 inner steps, modified Gram-Schmidt with the Kelley re-orthogonalization test,
 * Givens rotations, absolute tolerance on the preconditioned residual estimate
 * (the reference calls it at source/bem_problem.cc:826-827,853). */
static int gmres(const gm_op *o, const double *b, double *x, double tol, int max_steps,
                 int n_tmp, int *iters, double *last_res)
{
    const int n = o->n;
    const int m = n_tmp - 2;
    double *V = (double *)malloc(sizeof(double) * (size_t)n * (m + 1));
    double *p = (double *)malloc(sizeof(double) * n);
    double *Hh = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
    double *h = (double *)malloc(sizeof(double) * (m + 1));
    double *gamma = (double *)malloc(sizeof(double) * (m + 1));
    double *ci = (double *)malloc(sizeof(double) * m);
    double *si = (double *)malloc(sizeof(double) * m);
    double *y = (double *)malloc(sizeof(double) * m);
    int acc = 0, state = 0;
    int reorth = 0;
    double rho = 0;
    do {
        double *v0 = V;
        op_apply(o, x, p);
        for (int i = 0; i < n; ++i) p[i] = -p[i] + b[i];
        prec_apply(o, p, v0);
        rho = l2_norm(n, v0);
        *last_res = rho;
        state = sc_check(acc, rho, tol, max_steps);
        if (state != 0) break;
        gamma[0] = rho;
        for (int i = 0; i < n; ++i) v0[i] *= 1. / rho;
        int dim = 0;
        for (int inner = 0; inner < m && state == 0; ++inner) {
            ++acc;
            double *vv = V + (size_t)n * (inner + 1);
            op_apply(o, V + (size_t)n * inner, p);
            prec_apply(o, p, vv);
            dim = inner + 1;
            /* modified Gram-Schmidt */
            double norm_start = 0;
            if (!reorth && inner % 5 == 4) norm_start = l2_norm(n, vv);
            for (int i = 0; i < dim; ++i) {
                const double *vi = V + (size_t)n * i;
                h[i] = dot(n, vv, vi);
                for (int t = 0; t < n; ++t) vv[t] -= h[i] * vi[t];
            }
            double s = l2_norm(n, vv);
            if (!reorth && inner % 5 == 4) {
                if (!(s > 10. * norm_start * sqrt(__DBL_EPSILON__))) reorth = 1;
            }
            if (reorth) {
                for (int i = 0; i < dim; ++i) {
                    const double *vi = V + (size_t)n * i;
                    const double ht = dot(n, vv, vi);
                    h[i] += ht;
                    for (int t = 0; t < n; ++t) vv[t] -= ht * vi[t];
                }
                s = l2_norm(n, vv);
            }
            h[inner + 1] = s;
            for (int t = 0; t < n; ++t) vv[t] *= 1. / s;
            /* Givens rotations */
            for (int i = 0; i < inner; ++i) {
                const double ss = si[i], cc = ci[i], d0 = h[i];
                h[i] = cc * d0 + ss * h[i + 1];
                h[i + 1] = -ss * d0 + cc * h[i + 1];
            }
            const double r = 1. / sqrt(h[inner] * h[inner] + h[inner + 1] * h[inner + 1]);
            si[inner] = h[inner + 1] * r;
            ci[inner] = h[inner] * r;
            h[inner] = ci[inner] * h[inner] + si[inner] * h[inner + 1];
            gamma[inner + 1] = -si[inner] * gamma[inner];
            gamma[inner] *= ci[inner];
            for (int i = 0; i < dim; ++i) Hh[(size_t)i * m + inner] = h[i];
            rho = fabs(gamma[dim]);
            *last_res = rho;
            state = sc_check(acc, rho, tol, max_steps);
        }
        /* backward substitution H1 y = gamma, x += sum y_i v_i */
        for (int i = dim - 1; i >= 0; --i) {
            double s = gamma[i];
            for (int j = i + 1; j < dim; ++j) s -= Hh[(size_t)i * m + j] * y[j];
            y[i] = s / Hh[(size_t)i * m + i];
        }
        for (int i = 0; i < dim; ++i) {
            const double *vi = V + (size_t)n * i;
            for (int t = 0; t < n; ++t) x[t] += y[i] * vi[t];
        }
    } while (state == 0);
    *iters = acc;
    free(V);
    free(p);
    free(Hh);
    free(h);
    free(gamma);
    free(ci);
    free(si);
    free(y);
    return state == 1 ? 0 : 1;
}

/* BEMProblem::solve_system (source/bem_problem.cc:821-895), Direct branch, with the
 * constraint lines already computed (compute_constraints above). */
int orc_solve_system(int n, const double *G, const double *H, const double *surface,
                     const double *other, int n_lines, const int *line_rows, const int *line_ptr,
                     const int *ent_col, const double *ent_w, const double *inhom,
                     const double *tmp_rhs, double tol, int max_steps, int restart_vectors,
                     int precond, int band, double *phi, double *dphi_dn, double *alpha_out,
                     double *sol_out, int *iterations, double *last_residual)
{
    double *alpha = (double *)malloc(sizeof(double) * n);
    double *rhs = (double *)malloc(sizeof(double) * n);
    double *sol = (double *)calloc(n, sizeof(double));
    char *cons = (char *)calloc(n, 1);
    orc_compute_alpha(n, H, alpha);
    orc_compute_rhs(n, G, H, alpha, surface, other, tmp_rhs, rhs);
    for (int l = 0; l < n_lines; ++l) {
        rhs[line_rows[l]] = inhom[l]; /* ConstrainedOperator::distribute_rhs */
        cons[line_rows[l]] = 1;
    }
    gm_op o;
    o.n = n;
    o.G = G;
    o.H = H;
    o.alpha = alpha;
    o.surface = surface;
    o.other = other;
    o.n_lines = n_lines;
    o.line_rows = line_rows;
    o.line_ptr = line_ptr;
    o.ent_col = ent_col;
    o.ent_w = ent_w;
    o.precond = precond;
    o.lu.ab = NULL;
    o.lu.ipiv = NULL;
    int rc = 0;
    if (precond == 1) {
        if (band_precond_build(&o.lu, n, band, G, H, alpha, surface, cons) != 0) rc = -8;
    }
    if (rc == 0) {
        rc = gmres(&o, rhs, sol, tol, max_steps, restart_vectors, iterations, last_residual);
        if (rc == 0) {
            for (int i = 0; i < n; ++i) {
                if (surface[i] == 0)
                    phi[i] = sol[i];
                else
                    dphi_dn[i] = sol[i];
            }
        }
    }
    if (alpha_out) memcpy(alpha_out, alpha, sizeof(double) * n);
    if (sol_out) memcpy(sol_out, sol, sizeof(double) * n);
    free(o.lu.ab);
    free(o.lu.ipiv);
    free(alpha);
    free(rhs);
    free(sol);
    free(cons);
    return rc;
}

/* BEMProblem::residual (source/bem_problem.cc:903-961), Direct branch: constraints are
 * those computed from tmp = dphi_dn*other + phi*surface. */
void orc_residual(int n, const double *G, const double *H, const double *surface,
                  const double *other, int n_lines, const int *line_rows, const int *line_ptr,
                  const int *ent_col, const double *ent_w, const double *inhom,
                  const double *phi, const double *dphi_dn, double *res)
{
    double *alpha = (double *)malloc(sizeof(double) * n);
    double *tmp = (double *)malloc(sizeof(double) * n);
    double *rr = (double *)malloc(sizeof(double) * n);
    double *sol = (double *)malloc(sizeof(double) * n);
    orc_compute_alpha(n, H, alpha);
    for (int i = 0; i < n; ++i) tmp[i] = 0. + dphi_dn[i] * other[i] + phi[i] * surface[i];
    orc_compute_rhs(n, G, H, alpha, surface, other, tmp, rr);
    for (int l = 0; l < n_lines; ++l) rr[line_rows[l]] = inhom[l];
    for (int i = 0; i < n; ++i) rr[i] *= -1;
    for (int i = 0; i < n; ++i) sol[i] = 0. + dphi_dn[i] * surface[i] + phi[i] * other[i];
    orc_constrained_vmult(n, G, H, alpha, surface, other, n_lines, line_rows, line_ptr, ent_col,
                          ent_w, sol, res);
    for (int i = 0; i < n; ++i) res[i] += rr[i];
    free(alpha);
    free(tmp);
    free(rr);
    free(sol);
}

/* ------------------------------------------------------------------------------------
 * Backward of the assembly: out(d, :) = sum_i w_i dR_i/dx_d for the support point x_d of
 * every dof d, where R_i = (alpha_i phi_i + sum_j H_ij phi_j) - sum_j G_ij psi_j is the
 * BEM row of source/bem_problem.cc:621-717 (vmult + compute_rhs on the full vectors), with
 * alpha_i = -sum_j H_ij (compute_alpha, :594-618), so that
 *   R_i = sum_j H_ij (phi_j - phi_i) - sum_j G_ij psi_j.
 * Each (row i, cell c, point q) term of assemble_system (:106-590) is
 *   s = w_i/(4 pi) [ -(Phi_q - phi_i) (R.N) / r^3 - Psi_q J / r ],
 * R = y_q - x_i, N = n_q JxW_q = o w_q (a x b), J = JxW_q = w_q |a x b|, Phi_q / Psi_q the Q1
 * interpolants of phi / psi at q; its derivatives with respect to R, N and J are chained
 * through y = sum phi_k v_k, a = sum dxi_k v_k, b = sum deta_k v_k to the cell vertices, and
 * -dS/dR goes to x_i.  Regular and singular (QGaussOneOverR) points alike, with the same
 * rule selection as assemble_system.  Plain loops, no parallelism: small meshes only.
 * ---------------------------------------------------------------------------------- */
int orc_residual_vjp(int n, int nc, const int *cell_dofs, const signed char *orient,
                     const double *support, const int *dn_ptr, const int *dn_idx, int q_order,
                     int sing_order, const double *w, const double *phi, const double *psi,
                     double *out)
{
    if (n <= 0 || nc <= 0 || q_order <= 0 || q_order > 16 || sing_order <= 0 ||
        sing_order > 16)
        return -1;
    const int nq = q_order * q_order;
    const int ns = 2 * sing_order * sing_order;
    double *qpts = (double *)malloc(sizeof(double) * 2 * nq);
    double *qw = (double *)malloc(sizeof(double) * nq);
    double *spts = (double *)malloc(sizeof(double) * 4 * 2 * ns);
    double *sw = (double *)malloc(sizeof(double) * 4 * ns);
    if (!qpts || !qw || !spts || !sw) return -2;
    orc_qgauss_2d(q_order, qpts, qw);
    for (int k = 0; k < 4; ++k)
        orc_qgauss_one_over_r(sing_order, k, 1, spts + 2 * ns * k, sw + ns * k);
    memset(out, 0, sizeof(double) * 3 * (size_t)n);
    const double inv4pi = 1.0 / (4.0 * ORC_PI);
    for (int i = 0; i < n; ++i) {
        if (w[i] == 0.0) continue;
        const double *xi = support + 3 * i;
        const double wi = w[i] * inv4pi;
        for (int c = 0; c < nc; ++c) {
            const int *dofs = cell_dofs + 4 * c;
            int sing = -1;
            for (int j = 0; j < 4; ++j)
                if (in_set(dn_ptr, dn_idx, i, dofs[j])) {
                    sing = j;
                    break;
                }
            const double *pts = sing < 0 ? qpts : spts + 2 * ns * sing;
            const double *pw = sing < 0 ? qw : sw + ns * sing;
            const int np = sing < 0 ? nq : ns;
            double v[12];
            for (int k = 0; k < 4; ++k)
                for (int d = 0; d < 3; ++d) v[3 * k + d] = support[3 * dofs[k] + d];
            const double so = orient[c] < 0 ? -1.0 : 1.0;
            for (int q = 0; q < np; ++q) {
                double sh[4], dx[4], de[4];
                q1_shape(pts[2 * q], pts[2 * q + 1], sh, dx, de);
                double y[3] = {0, 0, 0}, a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
                double Phi = 0.0, Psi = 0.0;
                for (int k = 0; k < 4; ++k) {
                    for (int d = 0; d < 3; ++d) {
                        y[d] += sh[k] * v[3 * k + d];
                        a[d] += dx[k] * v[3 * k + d];
                        b[d] += de[k] * v[3 * k + d];
                    }
                    Phi += sh[k] * phi[dofs[k]];
                    Psi += sh[k] * psi[dofs[k]];
                }
                const double cx[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                                      a[0] * b[1] - a[1] * b[0]};
                const double cl = sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
                const double N[3] = {so * pw[q] * cx[0], so * pw[q] * cx[1], so * pw[q] * cx[2]};
                const double J = pw[q] * cl;
                const double R[3] = {y[0] - xi[0], y[1] - xi[1], y[2] - xi[2]};
                const double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
                const double ri = 1.0 / r, ri3 = ri * ri * ri, ri5 = ri3 * ri * ri;
                const double RN = R[0] * N[0] + R[1] * N[1] + R[2] * N[2];
                const double D = Phi - phi[i], E = Psi;
                double gR[3], gN[3];
                for (int d = 0; d < 3; ++d) {
                    gR[d] = wi * (-D * (N[d] * ri3 - 3.0 * RN * ri5 * R[d]) + E * J * ri3 * R[d]);
                    gN[d] = wi * (-D * ri3 * R[d]);
                }
                const double gJ = wi * (-E * ri);
                /* x_i */
                for (int d = 0; d < 3; ++d) out[3 * i + d] -= gR[d];
                /* N = so w (a x b): d/da = so w (b x gN), d/db = so w (gN x a);
                 * J = w |a x b|: d/da = w gJ (b x m), d/db = w gJ (m x a), m = (a x b)/|a x b| */
                const double m[3] = {cx[0] / cl, cx[1] / cl, cx[2] / cl};
                const double sN = so * pw[q], sJ = pw[q] * gJ;
                const double ga[3] = {
                    sN * (b[1] * gN[2] - b[2] * gN[1]) + sJ * (b[1] * m[2] - b[2] * m[1]),
                    sN * (b[2] * gN[0] - b[0] * gN[2]) + sJ * (b[2] * m[0] - b[0] * m[2]),
                    sN * (b[0] * gN[1] - b[1] * gN[0]) + sJ * (b[0] * m[1] - b[1] * m[0])};
                const double gb[3] = {
                    sN * (gN[1] * a[2] - gN[2] * a[1]) + sJ * (m[1] * a[2] - m[2] * a[1]),
                    sN * (gN[2] * a[0] - gN[0] * a[2]) + sJ * (m[2] * a[0] - m[0] * a[2]),
                    sN * (gN[0] * a[1] - gN[1] * a[0]) + sJ * (m[0] * a[1] - m[1] * a[0])};
                for (int k = 0; k < 4; ++k)
                    for (int d = 0; d < 3; ++d)
                        out[3 * dofs[k] + d] += sh[k] * gR[d] + dx[k] * ga[d] + de[k] * gb[d];
            }
        }
    }
    free(qpts);
    free(qw);
    free(spts);
    free(sw);
    return 0;
}

/* ------------------------------------------------------------------------------------
 * L2 projections onto the continuous Q1 space of the dofs (one mass matrix, three
 * components), solved directly (dense Cholesky; the reference uses SparseDirectUMFPACK):
 *   kind 0: ComputationalDomain::compute_normals (source/computational_domain.cc:1525-1620),
 *           rhs_i = sum_q phi_i n_q JxW_q, then every nodal vector normalised; without
 *           the normalisation NumericalTowingTank::compute_normals_at_nodes
 *           (source/numerical_towing_tank.cc:3358-3433);
 *   kind 1: BEMProblem::compute_surface_gradients (source/bem_problem.cc:1154-1290),
 *           rhs_i = sum_q phi_i grad_s(u)_q JxW_q with u = phi (the caller passes
 *           tmp_rhs .* surface_nodes), grad_s the FEValues gradient of the Q1 interpolant
 *           (J (J^T J)^-1 of the reference gradient).
 * ---------------------------------------------------------------------------------- */
static int l2_project(int n, int nc, const int *cell_dofs, const signed char *orient,
                      const double *support, int q_order, int kind, int normalize,
                      const double *u, double *out)
{
    const int nq = q_order * q_order;
    double *pts = (double *)malloc(sizeof(double) * 2 * nq);
    double *w = (double *)malloc(sizeof(double) * nq);
    double *M = (double *)calloc((size_t)n * n, sizeof(double));
    double *B = (double *)calloc((size_t)n * 3, sizeof(double));
    if (!pts || !w || !M || !B) return -2;
    orc_qgauss_2d(q_order, pts, w);
    for (int c = 0; c < nc; ++c) {
        const int *dofs = cell_dofs + 4 * c;
        double v[12];
        for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) v[3 * k + d] = support[3 * dofs[k] + d];
        for (int q = 0; q < nq; ++q) {
            double sh[4], dx[4], de[4];
            q1_shape(pts[2 * q], pts[2 * q + 1], sh, dx, de);
            double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
            for (int k = 0; k < 4; ++k)
                for (int d = 0; d < 3; ++d) {
                    a[d] += dx[k] * v[3 * k + d];
                    b[d] += de[k] * v[3 * k + d];
                }
            const double g00 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
            const double g01 = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
            const double g11 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
            const double det = g00 * g11 - g01 * g01;
            const double jxw = sqrt(det) * w[q];
            double f[3];
            if (kind == 0) {
                double nn[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                                a[0] * b[1] - a[1] * b[0]};
                const double nl = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
                for (int d = 0; d < 3; ++d) f[d] = nn[d] / nl * (orient[c] < 0 ? -1.0 : 1.0);
            } else {
                double ux = 0.0, ue = 0.0;
                for (int k = 0; k < 4; ++k) {
                    ux += dx[k] * u[dofs[k]];
                    ue += de[k] * u[dofs[k]];
                }
                const double ca = (g11 * ux - g01 * ue) / det, cb = (-g01 * ux + g00 * ue) / det;
                for (int d = 0; d < 3; ++d) f[d] = ca * a[d] + cb * b[d];
            }
            for (int i = 0; i < 4; ++i) {
                for (int j = 0; j < 4; ++j) M[(size_t)dofs[i] * n + dofs[j]] += sh[i] * sh[j] * jxw;
                for (int d = 0; d < 3; ++d) B[3 * (size_t)dofs[i] + d] += sh[i] * f[d] * jxw;
            }
        }
    }
    /* Cholesky M = L L^T in place (lower triangle) */
    for (int j = 0; j < n; ++j) {
        double s = M[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) s -= M[(size_t)j * n + k] * M[(size_t)j * n + k];
        if (s <= 0.0) {
            free(pts); free(w); free(M); free(B);
            return -3;
        }
        const double ljj = sqrt(s);
        M[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double t = M[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) t -= M[(size_t)i * n + k] * M[(size_t)j * n + k];
            M[(size_t)i * n + j] = t / ljj;
        }
    }
    for (int d = 0; d < 3; ++d) {
        for (int i = 0; i < n; ++i) {  /* L y = b */
            double t = B[3 * (size_t)i + d];
            for (int k = 0; k < i; ++k) t -= M[(size_t)i * n + k] * out[3 * (size_t)k + d];
            out[3 * (size_t)i + d] = t / M[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {  /* L^T x = y */
            double t = out[3 * (size_t)i + d];
            for (int k = i + 1; k < n; ++k) t -= M[(size_t)k * n + i] * out[3 * (size_t)k + d];
            out[3 * (size_t)i + d] = t / M[(size_t)i * n + i];
        }
    }
    if (kind == 0 && normalize)
        for (int i = 0; i < n; ++i) {
            double *o = out + 3 * (size_t)i;
            const double l = sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
            for (int d = 0; d < 3; ++d) o[d] /= l;
        }
    free(pts);
    free(w);
    free(M);
    free(B);
    return 0;
}

int orc_node_normals(int n, int nc, const int *cell_dofs, const signed char *orient,
                     const double *support, int q_order, int normalize, double *normals)
{
    return l2_project(n, nc, cell_dofs, orient, support, q_order, 0, normalize, NULL, normals);
}

int orc_surface_gradients(int n, int nc, const int *cell_dofs, const signed char *orient,
                          const double *support, int q_order, const double *phi,
                          double *grads)
{
    return l2_project(n, nc, cell_dofs, orient, support, q_order, 1, 0, phi, grads);
}
