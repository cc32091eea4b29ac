/*
 * bem_fma.c -- CPU restatement of WaveBEM's fast multipole matrix-vector product.
 *
 * TEST INFRASTRUCTURE / CPU BASELINE ONLY (see bem_oracle.h): nothing in wavebem_amd/
 * links or calls this code.
 *
 * Restates (paths relative to the WaveBEM source tree):
 *   source/computational_domain.cc:312-1453  ComputationalDomain::generate_octree_blocking
 *   source/octree_block.cc                   OctreeBlock (lists as sorted sets)
 *   source/bem_fma.cc:56-661                 BEMFMA::direct_integrals (near-field matrices)
 *   source/bem_fma.cc:663-765                BEMFMA::multipole_integrals
 *   source/bem_fma.cc:767-884                BEMFMA::generate_multipole_expansions
 *   source/bem_fma.cc:886-1087               BEMFMA::multipole_matr_vect_products
 *   source/multipole_expansion.cc            P2M (charge, normal dipole), M2M, M2P
 *   source/local_expansion.cc, include/local_expansion.h   M2L, L2L, L2P and their
 *                                            coefficient tables
 *   source/ass_leg_function.cc               normalised associated Legendre functions
 * The reference evaluates sqrt((n-m)!/(n+m)!) P_n^m(x) (Condon-Shortley phase) from
 * generated closed forms for n <= 20; here they come from the standard three-term
 * recurrence, equal to rounding.  Products: prodN = H phi and prodD = G dphi_dn of the
 * dense operator (bem_problem.cc:656-664), approximated by the multipole expansion of
 * truncation order p outside the near field.  FMA_preconditioner (bem_fma.cc:1090-1195)
 * is not restated.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "bem_oracle.h"

#define FMA_TOLL 0.001 /* source/computational_domain.cc:16 */
#define FMA_PMAX 10    /* coefficient tables of local_expansion.cc:7-10 */
#define FMA_LMAX 20    /* ass_leg_function.cc: n <= 20 */

/* ------------------------------------------------------------------ small containers */
typedef struct {
    int *a;
    int n, cap;
} ivec;

static void iv_push(ivec *v, int x)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 8;
        v->a = (int *)realloc(v->a, sizeof(int) * v->cap);
    }
    v->a[v->n++] = x;
}

static int cmp_int(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static void iv_sort_unique(ivec *v)
{
    if (v->n < 2) return;
    qsort(v->a, v->n, sizeof(int), cmp_int);
    int k = 1;
    for (int i = 1; i < v->n; ++i)
        if (v->a[i] != v->a[k - 1]) v->a[k++] = v->a[i];
    v->n = k;
}

static int iv_has(const ivec *v, int x) /* v sorted */
{
    int lo = 0, hi = v->n - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (v->a[mid] == x) return 1;
        if (v->a[mid] < x)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return 0;
}

static int lower_bound(const ivec *v, int x) /* first index with a >= x */
{
    int lo = 0, hi = v->n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v->a[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

static void iv_free(ivec *v)
{
    free(v->a);
    v->a = NULL;
    v->n = v->cap = 0;
}

/* ------------------------------------------------------------------ Legendre functions */
/* NF[n][m] = sqrt((n-m)!/(n+m)!): the normalisation of ass_leg_function.cc P_n_m */
static double NF[FMA_LMAX + 1][FMA_LMAX + 1];
/* A_n^m = (-1)^n / sqrt((n-m)! (n+m)!) (multipole_expansion.h A_n_m_Matrix) */
static double A_nm[FMA_LMAX + 1][FMA_LMAX + 1];
/* translation coefficients (local_expansion.h mExp_to_lExp_Coeff_Build,
 * lExp_to_lExp_Coeff_Build), [n][m][nn][mm + FMA_PMAX] */
static double C_M2L[FMA_PMAX + 1][FMA_PMAX + 1][FMA_PMAX + 1][2 * FMA_PMAX + 1];
static double C_L2L[FMA_PMAX + 1][FMA_PMAX + 1][FMA_PMAX + 1][2 * FMA_PMAX + 1];
static int tables_ready = 0;

/* real part of i^k for even k (every exponent used below is even) */
static double ipow_re(int k)
{
    const int r = ((k % 4) + 4) % 4;
    return r == 0 ? 1.0 : (r == 2 ? -1.0 : 0.0);
}

static void init_tables(void)
{
    if (tables_ready) return;
    for (int n = 0; n <= FMA_LMAX; ++n)
        for (int m = 0; m <= n; ++m) {
            double f1 = 1.0, f2 = 1.0, r = 1.0;
            for (int i = n - m; i > 0; --i) f1 *= i;
            for (int i = n + m; i > 0; --i) f2 *= i;
            A_nm[n][m] = pow(-1.0, (double)n) / sqrt(f1 * f2);
            for (int k = n - m + 1; k <= n + m; ++k) r *= k;
            NF[n][m] = 1.0 / sqrt(r);
        }
    for (int n = 0; n <= FMA_PMAX; ++n)
        for (int m = 0; m <= n; ++m)
            for (int nn = 0; nn <= FMA_PMAX; ++nn)
                for (int mm = -nn; mm <= nn; ++mm) {
                    double f = A_nm[nn][abs(mm)] / A_nm[n + nn][abs(m - mm)] * A_nm[n][abs(m)];
                    f *= ipow_re(abs(m - mm) - abs(m) - abs(mm)) / pow(-1.0, (double)nn);
                    C_M2L[n][m][nn][mm + FMA_PMAX] = f;
                    if (nn >= n && abs(mm - m) <= nn - n) {
                        double g = A_nm[nn - n][abs(mm - m)] / A_nm[nn][abs(mm)] * A_nm[n][abs(m)];
                        g *= ipow_re(abs(mm) - abs(mm - m) - abs(m)) * pow(-1.0, (double)(nn + n));
                        C_L2L[n][m][nn][mm + FMA_PMAX] = g;
                    }
                }
    tables_ready = 1;
}

/* P[n][m] = sqrt((n-m)!/(n+m)!) P_n^m(x) (Condon-Shortley phase) for n <= nmax, by the
 * three-term recurrence; optionally dS[n][m] = sqrt(1-x^2) d/dx of the same, from
 * (1 - x^2) dP_n^m/dx = (n + m) P_{n-1}^m - n x P_n^m.  Row stride FMA_LMAX + 1. */
#define LS (FMA_LMAX + 1)
static void legtab(int nmax, double x, double *P, double *dS)
{
    const double s = sqrt(1.0 - x * x);
    double U[LS * LS];
    double pmm = 1.0;
    for (int m = 0; m <= nmax; ++m) {
        if (m > 0) pmm *= -(2.0 * m - 1.0) * s;
        U[m * LS + m] = pmm;
        if (m + 1 <= nmax) U[(m + 1) * LS + m] = x * (2.0 * m + 1.0) * pmm;
        for (int k = m + 2; k <= nmax; ++k)
            U[k * LS + m] = ((2.0 * k - 1.0) * x * U[(k - 1) * LS + m] -
                             (k + m - 1.0) * U[(k - 2) * LS + m]) / (k - m);
    }
    for (int n = 0; n <= nmax; ++n)
        for (int m = 0; m <= n; ++m) {
            P[n * LS + m] = U[n * LS + m] * NF[n][m];
            if (dS) {
                const double pn1 = (n - 1 >= m) ? U[(n - 1) * LS + m] : 0.0;
                dS[n * LS + m] =
                    n == 0 ? 0.0 : ((n + m) * pn1 - n * x * U[n * LS + m]) / s * NF[n][m];
            }
        }
}

static int sgn(int x) { return (x > 0) - (x < 0); }

/* ------------------------------------------------------------------ expansions */
/* coefficients c[n(n+1)/2 + m], 0 <= m <= n <= p */
typedef struct {
    double complex *c;
    double ctr[3];
    int zero;
} Expansion;

static int ncoef(int p) { return (p + 1) * (p + 2) / 2; }
#define CI(n, m) ((n) * ((n) + 1) / 2 + (m))

static void sph(const double *rel, double *rho, double *cosa, double *beta)
{
    *rho = sqrt(rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2]);
    *cosa = rel[2] / *rho;
    *beta = atan2(rel[1], rel[0]);
}

/* e^{i k beta} for k in [-K, K], stored at E[k + K] */
static void phases(int K, double beta, double complex *E)
{
    for (int k = -K; k <= K; ++k) E[k + K] = cos(k * beta) + I * sin(k * beta);
}

/* MultipoleExpansion::Add(strength, point) */
static void mp_add_point(Expansion *M, int p, double strength, const double *y)
{
    if (strength == 0.0) return;
    M->zero = 0;
    const double rel[3] = {y[0] - M->ctr[0], y[1] - M->ctr[1], y[2] - M->ctr[2]};
    double rho, ca, beta, P[LS * LS];
    double complex E[2 * FMA_PMAX + 1];
    sph(rel, &rho, &ca, &beta);
    legtab(p, ca, P, NULL);
    phases(p, beta, E);
    double rn = 1.0;
    for (int n = 0; n <= p; ++n, rn *= rho)
        for (int m = 0; m <= n; ++m) M->c[CI(n, m)] += E[p - m] * (P[n * LS + m] * rn * strength);
}

/* MultipoleExpansion::AddNormDer(strength, point, normal) */
static void mp_add_normder(Expansion *M, int p, double strength, const double *y,
                           const double *normal)
{
    if (strength == 0.0) return;
    M->zero = 0;
    const double rel[3] = {y[0] - M->ctr[0], y[1] - M->ctr[1], y[2] - M->ctr[2]};
    const double nl = sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
    const double nv[3] = {normal[0] / nl, normal[1] / nl, normal[2] / nl};
    double rho, ca, beta, P[LS * LS], dS[LS * LS];
    double complex E[2 * FMA_PMAX + 1];
    sph(rel, &rho, &ca, &beta);
    const double drho = (rel[0] / rho) * nv[0] + (rel[1] / rho) * nv[1] + (rel[2] / rho) * nv[2];
    const double rxy = rel[0] * rel[0] + rel[1] * rel[1];
    const double dbeta = (-rel[1] / rxy) * nv[0] + (rel[0] / rxy) * nv[1];
    const double sa = sqrt(1.0 - ca * ca);
    const double dalpha = (ca * cos(beta) / rho) * nv[0] + (ca * sin(beta) / rho) * nv[1] +
                          (-sa / rho) * nv[2];
    legtab(p, ca, P, dS);
    phases(p, beta, E);
    for (int n = 0; n <= p; ++n) {
        const double rn = pow(rho, (double)n), rn1 = n * pow(rho, n - 1.0);
        for (int m = 0; m <= n; ++m) {
            const double Pn = P[n * LS + m];
            const double complex w =
                (rn1 * Pn * drho - rn * dS[n * LS + m] * dalpha) + I * (-(double)m * rn * Pn * dbeta);
            M->c[CI(n, m)] += E[p - m] * w * strength;
        }
    }
}

/* MultipoleExpansion::Add(multipole, sol) */
static void mp_add_scaled(Expansion *M, int p, const Expansion *E, double sol)
{
    if (E->zero || sol == 0.0) return;
    M->zero = 0;
    const int nc = ncoef(p);
    for (int k = 0; k < nc; ++k) M->c[k] += E->c[k] * sol;
}

/* MultipoleExpansion::Add(child): translation of a child multipole to this centre */
static void mp_translate(Expansion *M, int p, const Expansion *C)
{
    if (C->zero) return;
    M->zero = 0;
    const double rel[3] = {C->ctr[0] - M->ctr[0], C->ctr[1] - M->ctr[1], C->ctr[2] - M->ctr[2]};
    const double dist = sqrt(rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2]);
    if (!(dist > 1e-7)) {
        for (int k = 0; k < ncoef(p); ++k) M->c[k] += C->c[k];
        return;
    }
    double rho, ca, beta, P[LS * LS], rp[FMA_PMAX + 1];
    double complex E[2 * FMA_PMAX + 1];
    sph(rel, &rho, &ca, &beta);
    legtab(p, ca, P, NULL);
    phases(p, beta, E);
    rp[0] = 1.0;
    for (int k = 1; k <= p; ++k) rp[k] = rp[k - 1] * rho;
    for (int n = 0; n <= p; ++n)
        for (int m = 0; m <= n; ++m) {
            double complex z = 0.0;
            for (int nn = 0; nn <= n; ++nn)
                for (int mm = -nn; mm <= nn; ++mm) {
                    if (abs(m - mm) > n - nn) continue;
                    const double complex o = C->c[CI(n - nn, abs(m - mm))];
                    const double complex a = creal(o) + I * (sgn(m - mm) * cimag(o));
                    double f = P[nn * LS + abs(mm)] * rp[nn] * A_nm[nn][abs(mm)] *
                               A_nm[n - nn][abs(m - mm)] / A_nm[n][m];
                    f *= ipow_re(abs(m) - abs(mm) - abs(m - mm));
                    z += f * (a * E[p - mm]);
                }
            M->c[CI(n, m)] += z;
        }
}

/* MultipoleExpansion::Evaluate */
static double mp_eval(const Expansion *M, int p, const double *x)
{
    if (M->zero) return 0.0;
    const double rel[3] = {x[0] - M->ctr[0], x[1] - M->ctr[1], x[2] - M->ctr[2]};
    double rho, ca, beta, P[LS * LS];
    double complex E[2 * FMA_PMAX + 1];
    sph(rel, &rho, &ca, &beta);
    legtab(p, ca, P, NULL);
    phases(p, beta, E);
    double complex v = 0.0;
    const double ir = 1.0 / rho;
    double rf = ir;
    for (int n = 0; n <= p; ++n, rf *= ir) {
        v += M->c[CI(n, 0)] * (P[n * LS] * rf);
        for (int m = 1; m <= n; ++m) v += M->c[CI(n, m)] * E[p + m] * 2.0 * (P[n * LS + m] * rf);
    }
    return creal(v);
}

/* LocalExpansion::Add(multipole): M2L */
static void lc_add_mp(Expansion *L, int p, const Expansion *M)
{
    if (M->zero) return;
    const double rel[3] = {M->ctr[0] - L->ctr[0], M->ctr[1] - L->ctr[1], M->ctr[2] - L->ctr[2]};
    double rho, ca, beta, P[LS * LS], rf[2 * FMA_PMAX + 2];
    double complex E[4 * FMA_PMAX + 1];
    sph(rel, &rho, &ca, &beta);
    legtab(2 * p, ca, P, NULL);
    phases(2 * p, beta, E);
    const double ir = 1.0 / rho;
    rf[0] = 1.0;
    for (int k = 1; k <= 2 * p + 1; ++k) rf[k] = rf[k - 1] * ir;
    for (int n = 0; n <= p; ++n)
        for (int m = 0; m <= n; ++m) {
            double complex z = 0.0;
            for (int nn = 0; nn <= p; ++nn) {
                const double r = rf[n + nn + 1];
                const double *cf = C_M2L[n][m][nn] + FMA_PMAX;
                for (int mm = -nn; mm < 0; ++mm) {
                    const double complex a = conj(M->c[CI(nn, -mm)]);
                    const double f = P[(nn + n) * LS + abs(mm - m)] * r * cf[mm];
                    z += a * E[2 * p + mm - m] * f;
                }
                for (int mm = 0; mm <= nn; ++mm) {
                    const double complex a = M->c[CI(nn, mm)];
                    const double f = P[(nn + n) * LS + abs(mm - m)] * r * cf[mm];
                    z += a * E[2 * p + mm - m] * f;
                }
            }
            L->c[CI(n, m)] += z;
        }
    L->zero = 0;
}

/* LocalExpansion::Add(local): L2L (translation of the parent's local expansion) */
static void lc_translate(Expansion *L, int p, const Expansion *Pa)
{
    if (Pa->zero) return;
    const double rel[3] = {Pa->ctr[0] - L->ctr[0], Pa->ctr[1] - L->ctr[1], Pa->ctr[2] - L->ctr[2]};
    const double dist = sqrt(rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2]);
    if (dist > 1e-7) {
        double rho, ca, beta, P[LS * LS], rp[FMA_PMAX + 1];
        double complex E[4 * FMA_PMAX + 1];
        sph(rel, &rho, &ca, &beta);
        legtab(p, ca, P, NULL);
        phases(2 * p, beta, E);
        rp[0] = 1.0;
        for (int k = 1; k <= p; ++k) rp[k] = rp[k - 1] * rho;
        for (int n = 0; n <= p; ++n)
            for (int m = 0; m <= n; ++m) {
                double complex z = 0.0;
                for (int nn = n; nn <= p; ++nn) {
                    const double r = rp[nn - n];
                    const double *cf = C_L2L[n][m][nn] + FMA_PMAX;
                    for (int mm = -nn; mm <= nn; ++mm) {
                        if (abs(mm - m) > nn - n) continue;
                        const double complex o = Pa->c[CI(nn, abs(mm))];
                        const double complex a = creal(o) + I * (sgn(mm) * cimag(o));
                        const double f = P[(nn - n) * LS + abs(mm - m)] * r * cf[mm];
                        z += a * E[2 * p + mm - m] * f;
                    }
                }
                L->c[CI(n, m)] += z;
            }
    } else {
        for (int k = 0; k < ncoef(p); ++k) L->c[k] += Pa->c[k];
    }
    L->zero = 0;
}

/* LocalExpansion::Evaluate */
static double lc_eval(const Expansion *L, int p, const double *x)
{
    if (L->zero) return 0.0;
    const double rel[3] = {x[0] - L->ctr[0], x[1] - L->ctr[1], x[2] - L->ctr[2]};
    double rho, ca, beta, P[LS * LS];
    double complex E[2 * FMA_PMAX + 1];
    sph(rel, &rho, &ca, &beta);
    legtab(p, ca, P, NULL);
    phases(p, beta, E);
    double complex v = 0.0;
    double rn = 1.0;
    for (int n = 0; n <= p; ++n, rn *= rho) {
        v += L->c[CI(n, 0)] * (P[n * LS] * rn);
        for (int m = 1; m <= n; ++m) v += L->c[CI(n, m)] * (E[p + m] * 2.0 * (P[n * LS + m] * rn));
    }
    return creal(v);
}

/* ------------------------------------------------------------------ octree */
typedef struct {
    int level, parent, nchild, child[8];
    double pmin[3], delta;
    ivec nodes; /* dofs, insertion order */
    ivec qp;    /* global quadrature point ids cell*nq + q, ascending */
    int nsub;   /* sublevels of the near-neighbour / interaction lists */
    ivec *nn, *il, *nil;
} Block;

/* near-field sparse rows */
typedef struct {
    int *col;
    double *h, *g;
    int n, cap;
} SpRow;

struct orc_fma {
    int n, nc, nq, p, L;
    const int *cd;
    int *dn_ptr, *dn_idx;
    double *X;                     /* support points */
    double *qx, *qn, *qj, *qs;     /* per quad point: point, normal, JxW, 4 shape values */
    int *cell_ptr, *cell_items;    /* dof -> cells (dof_to_elems) */
    Block *B;
    int nb;
    int *start, *end;              /* per level */
    ivec childless;
    ivec *filled;                  /* dofs_filled_blocks per level */
    int *qp2b;                     /* [nqp][L+1] */
    SpRow *rows;
    /* element multipole expansions of the leaves: entries (leaf, cell) */
    int *el_leaf_ptr;              /* per block: first entry (leaves only), size nb + 1 */
    int *el_cell;
    double complex *el_c;          /* [entry][4 dofs][2 kernels][ncoef] */
    Expansion *M1, *M2, *L1, *L2;  /* per block */
    double complex *mem;
    /* quadrature for singular cells */
    int sing_order;
    const signed char *orient;
    const double *support;
};

static void add_sp(SpRow *r, int col, double h, double g)
{
    for (int k = 0; k < r->n; ++k)
        if (r->col[k] == col) {
            r->h[k] += h;
            r->g[k] += g;
            return;
        }
    if (r->n == r->cap) {
        r->cap = r->cap ? 2 * r->cap : 16;
        r->col = (int *)realloc(r->col, sizeof(int) * r->cap);
        r->h = (double *)realloc(r->h, sizeof(double) * r->cap);
        r->g = (double *)realloc(r->g, sizeof(double) * r->cap);
    }
    r->col[r->n] = col;
    r->h[r->n] = h;
    r->g[r->n] = g;
    r->n++;
}

static void new_block(Block *b, int level, int parent, const double *pmin, double delta)
{
    memset(b, 0, sizeof(*b));
    b->level = level;
    b->parent = parent;
    b->pmin[0] = pmin[0];
    b->pmin[1] = pmin[1];
    b->pmin[2] = pmin[2];
    b->delta = delta;
    b->nsub = 1;
}

static void alloc_lists(Block *b, int nsub)
{
    for (int s = 0; s < b->nsub && b->nn; ++s) {
        iv_free(&b->nn[s]);
        iv_free(&b->il[s]);
        iv_free(&b->nil[s]);
    }
    free(b->nn);
    free(b->il);
    free(b->nil);
    b->nsub = nsub;
    b->nn = (ivec *)calloc(nsub, sizeof(ivec));
    b->il = (ivec *)calloc(nsub, sizeof(ivec));
    b->nil = (ivec *)calloc(nsub, sizeof(ivec));
}

/* child index for a point (computational_domain.cc:530-640) */
static int child_of(const Block *par, double delta, const double *x)
{
    const int up_z = !(x[2] <= par->pmin[2] + delta);
    const int up_y = !(x[1] <= par->pmin[1] + delta);
    const int up_x = !(x[0] <= par->pmin[0] + delta);
    static const int tab[2][2][2] = {{{0, 1}, {3, 2}}, {{4, 5}, {7, 6}}};
    return tab[up_z][up_y][up_x];
}

/* the face/edge/corner contact test of computational_domain.cc:1013-1046 */
static int adjacent(const Block *a, const Block *b)
{
    double pmin1[3], pmax1[3], pmin2[3], pmax2[3];
    for (int d = 0; d < 3; ++d) {
        pmin1[d] = a->pmin[d];
        pmax1[d] = 2.0 * (a->delta / 2.0) + a->pmin[d];
        pmin2[d] = b->pmin[d];
        pmax2[d] = 2.0 * (b->delta / 2.0) + b->pmin[d];
    }
    const double T = FMA_TOLL;
#define OVL(d) ((pmin1[d] - T <= pmax2[d]) && (pmax1[d] + T >= pmin2[d]))
#define TOUCH(d) ((fabs(pmin1[d] - pmax2[d]) <= T) || (fabs(pmax1[d] - pmin2[d]) <= T))
    if (TOUCH(0) && OVL(1) && OVL(2)) return 1;
    if (TOUCH(1) && OVL(0) && OVL(2)) return 1;
    if (TOUCH(2) && OVL(1) && OVL(0)) return 1;
#undef OVL
#undef TOUCH
    return 0;
}

static int dset_size(const orc_fma *f, int i) { return f->dn_ptr[i + 1] - f->dn_ptr[i]; }

static void build_octree(orc_fma *f)
{
    const int n = f->n, nq = f->nq, L = f->L;
    const int nqp = f->nc * nq;
    double maxc = 0.0;
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) maxc = fmax(maxc, fabs(f->X[3 * i + d]));
    const int maxb = L * (nqp + n) + 2;
    f->B = (Block *)calloc(maxb, sizeof(Block));
    f->start = (int *)calloc(L + 1, sizeof(int));
    f->end = (int *)calloc(L + 1, sizeof(int));
    f->filled = (ivec *)calloc(L + 1, sizeof(ivec));
    f->qp2b = (int *)malloc(sizeof(int) * (size_t)nqp * (L + 1));
    ivec *parents = (ivec *)calloc(L + 1, sizeof(ivec));
    double pmin[3] = {-1.1 * maxc, -1.1 * maxc, -1.1 * maxc};
    double delta = 2.2 * maxc;
    Block *root = &f->B[0];
    new_block(root, 0, 0, pmin, delta);
    for (int q = 0; q < nqp; ++q) {
        iv_push(&root->qp, q);
        f->qp2b[(size_t)q * (L + 1)] = 0;
    }
    for (int i = 0; i < n; ++i) iv_push(&root->nodes, i);
    int count = 0;
    iv_push(&parents[0], 0);
    for (int level = 1; level <= L; ++level) {
        delta /= 2.0;
        for (int k = 0; k < parents[level - 1].n; ++k) {
            const int jj = parents[level - 1].a[k];
            Block ch[8];
            const double off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                      {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
            for (int c = 0; c < 8; ++c) {
                double pm[3];
                for (int d = 0; d < 3; ++d) pm[d] = f->B[jj].pmin[d] + (off[c][d] ? delta : 0.0);
                new_block(&ch[c], level, jj, pm, delta);
            }
            const Block *par = &f->B[jj];
            for (int i = 0; i < par->nodes.n; ++i) {
                const int d = par->nodes.a[i];
                iv_push(&ch[child_of(par, delta, f->X + 3 * d)].nodes, d);
            }
            for (int i = 0; i < par->qp.n; ++i) {
                const int q = par->qp.a[i];
                iv_push(&ch[child_of(par, delta, f->qx + 3 * (size_t)q)].qp, q);
            }
            for (int c = 0; c < 8; ++c) {
                if (ch[c].nodes.n + ch[c].qp.n > 0) {
                    ++count;
                    f->B[count] = ch[c];
                    Block *pb = &f->B[jj];
                    pb->child[pb->nchild++] = count;
                    for (int i = 0; i < ch[c].qp.n; ++i)
                        f->qp2b[(size_t)ch[c].qp.a[i] * (L + 1) + level] = count;
                } else {
                    iv_free(&ch[c].nodes);
                    iv_free(&ch[c].qp);
                }
            }
        }
        f->start[level] = f->end[level - 1] + 1;
        f->end[level] = count;
        for (int jj = f->start[level]; jj <= f->end[level]; ++jj) {
            Block *b = &f->B[jj];
            int ndouble = 0;
            for (int i = 0; i < b->nodes.n; ++i) ndouble += dset_size(f, b->nodes.a[i]) - 1;
            if (b->nodes.n - ndouble < 2) {
                iv_push(&f->childless, jj);
                for (int i = 0; i < b->qp.n; ++i)
                    for (int j = level + 1; j <= L; ++j)
                        f->qp2b[(size_t)b->qp.a[i] * (L + 1) + j] = jj;
            } else {
                iv_push(&parents[level], jj);
            }
            if (b->nodes.n > 0) iv_push(&f->filled[level], jj);
        }
    }
    for (int k = 0; k < parents[L].n; ++k) iv_push(&f->childless, parents[L].a[k]);
    f->nb = count + 1;
    for (int l = 0; l <= L; ++l) iv_free(&parents[l]);
    free(parents);

    /* near neighbours (computational_domain.cc:934-1190) */
    for (int ii = f->start[1]; ii <= f->end[1]; ++ii) {
        alloc_lists(&f->B[ii], 1);
        for (int jj = f->start[1]; jj <= f->end[1]; ++jj) iv_push(&f->B[ii].nn[0], jj);
    }
    for (int level = 2; level <= L; ++level)
        for (int kk = f->start[level]; kk <= f->end[level]; ++kk) {
            Block *b1 = &f->B[kk];
            if (b1->nodes.n == 0) continue; /* the lists of node-free blocks are never read */
            const int nsub = b1->nchild == 0 ? L - level + 1 : 1;
            alloc_lists(b1, nsub);
            iv_push(&b1->nn[0], kk);
            const ivec *pn = &f->B[b1->parent].nn[0];
            for (int t = 0; t < pn->n; ++t) {
                const Block *b2 = &f->B[pn->a[t]];
                if (b2->nchild == 0 && adjacent(b1, b2)) iv_push(&b1->nn[0], pn->a[t]);
                for (int c = 0; c < b2->nchild; ++c)
                    if (adjacent(b1, &f->B[b2->child[c]])) iv_push(&b1->nn[0], b2->child[c]);
            }
            iv_sort_unique(&b1->nn[0]);
            for (int s = 1; s < nsub; ++s) {
                const ivec *up = &b1->nn[s - 1];
                for (int t = 0; t < up->n; ++t) {
                    const Block *b2 = &f->B[up->a[t]];
                    if (b2->nchild == 0) iv_push(&b1->nn[s], up->a[t]);
                    for (int c = 0; c < b2->nchild; ++c)
                        if (adjacent(b1, &f->B[b2->child[c]])) iv_push(&b1->nn[s], b2->child[c]);
                }
                iv_sort_unique(&b1->nn[s]);
            }
        }

    /* interaction lists (computational_domain.cc:1227-1380) */
    for (int ii = f->start[1]; ii <= f->end[1]; ++ii) {
        for (int jj = f->start[1]; jj <= f->end[1]; ++jj) iv_push(&f->B[ii].il[0], jj);
        iv_sort_unique(&f->B[ii].il[0]);
    }
    for (int level = 2; level <= L; ++level)
        for (int kk = f->start[level]; kk <= f->end[level]; ++kk) {
            Block *b1 = &f->B[kk];
            if (b1->nodes.n == 0) continue;
            for (int s = 0; s < b1->nsub; ++s) {
                ivec *il = &b1->il[s];
                for (int t = 0; t < b1->nn[s].n; ++t) iv_push(il, b1->nn[s].a[t]);
                for (int i = 0; i < b1->nodes.n; ++i) {
                    const int nd = b1->nodes.a[i];
                    for (int p = f->dn_ptr[nd]; p < f->dn_ptr[nd + 1]; ++p) {
                        const int dd = f->dn_idx[p];
                        for (int e = f->cell_ptr[dd]; e < f->cell_ptr[dd + 1]; ++e) {
                            const int cell = f->cell_items[e];
                            for (int j = 0; j < nq; ++j)
                                iv_push(il, f->qp2b[((size_t)cell * nq + j) * (L + 1) + level + s]);
                        }
                    }
                }
                iv_sort_unique(il);
            }
            for (int s = 0; s < b1->nsub; ++s) {
                const ivec *il = &b1->il[s];
                const ivec *pil = s == 0 ? &f->B[b1->parent].il[0] : &b1->il[s - 1];
                for (int t = 0; t < pil->n; ++t) {
                    const Block *b2 = &f->B[pil->a[t]];
                    if (b2->nchild == 0) {
                        if (!iv_has(il, pil->a[t])) iv_push(&b1->nil[s], pil->a[t]);
                    } else {
                        for (int c = 0; c < b2->nchild; ++c)
                            if (!iv_has(il, b2->child[c])) iv_push(&b1->nil[s], b2->child[c]);
                    }
                }
                iv_sort_unique(&b1->nil[s]);
            }
        }
}

/* ------------------------------------------------------------------ near field */
static void row_regular(orc_fma *f, int node, int cell, const int *qs, int nqs, double *lh,
                        double *lg)
{
    const double *x = f->X + 3 * node;
    for (int k = 0; k < nqs; ++k) {
        const size_t q = (size_t)cell * f->nq + qs[k];
        const double R[3] = {f->qx[3 * q] + -1.0 * x[0], f->qx[3 * q + 1] + -1.0 * x[1],
                             f->qx[3 * q + 2] + -1.0 * x[2]};
        const double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
        const double s = 1. / (r * 4 * M_PI);
        const double den = -4 * M_PI * (r * r) * r;
        const double D[3] = {R[0] / den, R[1] / den, R[2] / den};
        const double *nn = f->qn + 3 * q;
        const double Dn = D[0] * nn[0] + D[1] * nn[1] + D[2] * nn[2];
        for (int j = 0; j < 4; ++j) {
            lh[j] += Dn * f->qs[4 * q + j] * f->qj[q];
            lg[j] += s * f->qs[4 * q + j] * f->qj[q];
        }
    }
}

static void row_singular(orc_fma *f, int node, int cell, int vtx, double *lh, double *lg)
{
    const int so = f->sing_order, ns = 2 * so * so;
    double *pts = (double *)malloc(sizeof(double) * 2 * ns), *w = (double *)malloc(sizeof(double) * ns);
    double *qp = (double *)malloc(sizeof(double) * 3 * ns), *nr = (double *)malloc(sizeof(double) * 3 * ns);
    double *jxw = (double *)malloc(sizeof(double) * ns), *sh = (double *)malloc(sizeof(double) * 4 * ns);
    orc_qgauss_one_over_r(so, vtx, 1, pts, w);
    double v12[12];
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) v12[3 * k + d] = f->support[3 * f->cd[4 * cell + k] + d];
    orc_fe_values(v12, f->orient[cell], ns, pts, w, qp, nr, jxw, sh);
    const double *x = f->X + 3 * node;
    for (int q = 0; q < ns; ++q) {
        const double R[3] = {qp[3 * q] + -1.0 * x[0], qp[3 * q + 1] + -1.0 * x[1],
                             qp[3 * q + 2] + -1.0 * x[2]};
        const double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
        const double s = 1. / (r * 4 * M_PI);
        const double den = -4 * M_PI * (r * r) * r;
        const double D[3] = {R[0] / den, R[1] / den, R[2] / den};
        const double Dn = D[0] * nr[3 * q] + D[1] * nr[3 * q + 1] + D[2] * nr[3 * q + 2];
        for (int j = 0; j < 4; ++j) {
            lh[j] += Dn * sh[4 * q + j] * jxw[q];
            lg[j] += s * sh[4 * q + j] * jxw[q];
        }
    }
    free(pts);
    free(w);
    free(qp);
    free(nr);
    free(jxw);
    free(sh);
}

/* (cell, sorted q list) groups of the union of the blocks' quadrature points */
static int collect_points(const orc_fma *f, const int *blocks, int nblocks, int *buf)
{
    int m = 0;
    for (int t = 0; t < nblocks; ++t) {
        const ivec *qp = &f->B[blocks[t]].qp;
        for (int i = 0; i < qp->n; ++i) buf[m++] = qp->a[i];
    }
    qsort(buf, m, sizeof(int), cmp_int);
    int k = 0;
    for (int i = 0; i < m; ++i)
        if (k == 0 || buf[i] != buf[k - 1]) buf[k++] = buf[i];
    return k;
}

static void direct_integrals(orc_fma *f)
{
    const int nq = f->nq, L = f->L;
    f->rows = (SpRow *)calloc(f->n, sizeof(SpRow));
    int *buf = (int *)malloc(sizeof(int) * (size_t)f->nc * nq);
    int *qs = (int *)malloc(sizeof(int) * nq);
    /* childless blocks: their interaction list at the finest sublevel (bem_fma.cc:263-419) */
    for (int k = 0; k < f->childless.n; ++k) {
        const Block *b1 = &f->B[f->childless.a[k]];
        if (b1->nodes.n == 0) continue;
        const ivec *il = &b1->il[b1->nsub - 1];
        const int m = collect_points(f, il->a, il->n, buf);
        for (int i = 0; i < b1->nodes.n; ++i) {
            const int node = b1->nodes.a[i];
            for (int s = 0; s < m;) {
                const int cell = buf[s] / nq;
                int e = s, nqs = 0;
                while (e < m && buf[e] / nq == cell) qs[nqs++] = buf[e++] % nq;
                double lh[4] = {0, 0, 0, 0}, lg[4] = {0, 0, 0, 0};
                int vtx = -1;
                for (int j = 0; j < 4 && vtx < 0; ++j)
                    for (int p = f->dn_ptr[node]; p < f->dn_ptr[node + 1]; ++p)
                        if (f->dn_idx[p] == f->cd[4 * cell + j]) {
                            vtx = j;
                            break;
                        }
                if (vtx < 0)
                    row_regular(f, node, cell, qs, nqs, lh, lg);
                else
                    row_singular(f, node, cell, vtx, lh, lg);
                for (int j = 0; j < 4; ++j) add_sp(&f->rows[node], f->cd[4 * cell + j], lh[j], lg[j]);
                s = e;
            }
        }
    }
    /* childless blocks of coarser levels in the non-interaction lists (bem_fma.cc:522-640) */
    int *bl = (int *)malloc(sizeof(int) * (size_t)f->nb);
    for (int level = 1; level <= L; ++level) {
        const int sb = f->start[level];
        for (int k = 0; k < f->filled[level].n; ++k) {
            const Block *b1 = &f->B[f->filled[level].a[k]];
            int nbl = 0;
            for (int s = 0; s < b1->nsub && b1->nil; ++s) {
                const ivec *nil = &b1->nil[s];
                const int hi = lower_bound(nil, sb);
                for (int t = 0; t < hi; ++t) bl[nbl++] = nil->a[t];
            }
            if (nbl == 0) continue;
            const int m = collect_points(f, bl, nbl, buf);
            for (int i = 0; i < b1->nodes.n; ++i) {
                const int node = b1->nodes.a[i];
                for (int s = 0; s < m;) {
                    const int cell = buf[s] / nq;
                    int e = s, nqs = 0;
                    while (e < m && buf[e] / nq == cell) qs[nqs++] = buf[e++] % nq;
                    double lh[4] = {0, 0, 0, 0}, lg[4] = {0, 0, 0, 0};
                    row_regular(f, node, cell, qs, nqs, lh, lg);
                    for (int j = 0; j < 4; ++j)
                        add_sp(&f->rows[node], f->cd[4 * cell + j], lh[j], lg[j]);
                    s = e;
                }
            }
        }
    }
    free(bl);
    free(buf);
    free(qs);
}

/* ------------------------------------------------------------------ multipole integrals */
static void multipole_integrals(orc_fma *f)
{
    const int nq = f->nq, p = f->p, K = ncoef(p);
    f->el_leaf_ptr = (int *)calloc(f->nb + 1, sizeof(int));
    /* count (leaf, cell) entries */
    int *is_leaf = (int *)calloc(f->nb, sizeof(int));
    for (int k = 0; k < f->childless.n; ++k) is_leaf[f->childless.a[k]] = 1;
    long total = 0;
    for (int b = 0; b < f->nb; ++b) {
        f->el_leaf_ptr[b] = (int)total;
        if (!is_leaf[b]) continue;
        const ivec *qp = &f->B[b].qp;
        for (int i = 0; i < qp->n; ++i)
            if (i == 0 || qp->a[i] / nq != qp->a[i - 1] / nq) total++;
    }
    f->el_leaf_ptr[f->nb] = (int)total;
    f->el_cell = (int *)malloc(sizeof(int) * (total ? total : 1));
    f->el_c = (double complex *)calloc((size_t)(total ? total : 1) * 8 * K, sizeof(double complex));
    for (int b = 0; b < f->nb; ++b) {
        if (!is_leaf[b]) continue;
        const Block *blk = &f->B[b];
        double ctr[3];
        for (int d = 0; d < 3; ++d) ctr[d] = blk->pmin[d] + blk->delta / 2.;
        int e = f->el_leaf_ptr[b] - 1;
        for (int i = 0; i < blk->qp.n; ++i) {
            const int cell = blk->qp.a[i] / nq, q = blk->qp.a[i] % nq;
            if (i == 0 || blk->qp.a[i - 1] / nq != cell) {
                ++e;
                f->el_cell[e] = cell;
            }
            const size_t g = (size_t)cell * nq + q;
            for (int j = 0; j < 4; ++j) {
                Expansion e1 = {f->el_c + ((size_t)e * 8 + 2 * j) * K, {ctr[0], ctr[1], ctr[2]}, 1};
                Expansion e2 = {f->el_c + ((size_t)e * 8 + 2 * j + 1) * K, {ctr[0], ctr[1], ctr[2]}, 1};
                const double st = f->qs[4 * g + j] * f->qj[g] / 4 / M_PI;
                mp_add_normder(&e1, p, st, f->qx + 3 * g, f->qn + 3 * g);
                mp_add_point(&e2, p, st, f->qx + 3 * g);
            }
        }
    }
    free(is_leaf);
}

/* element expansions are zero only when all their strengths were 0 */
static int el_zero(const double complex *c, int K)
{
    for (int k = 0; k < K; ++k)
        if (c[k] != 0.0) return 0;
    return 1;
}

/* ------------------------------------------------------------------ public API */
orc_fma *orc_fma_create(int n, int nc, const int *cell_dofs, const signed char *orient,
                        const double *support, const int *dn_ptr, const int *dn_idx,
                        int q_order, int sing_order, int trunc_order, int num_levels)
{
    if (n <= 0 || nc <= 0 || trunc_order < 0 || trunc_order > FMA_PMAX || num_levels < 1 ||
        q_order < 1)
        return NULL;
    init_tables();
    orc_fma *f = (orc_fma *)calloc(1, sizeof(orc_fma));
    f->n = n;
    f->nc = nc;
    f->nq = q_order * q_order;
    f->p = trunc_order;
    f->L = num_levels;
    f->cd = cell_dofs;
    f->orient = orient;
    f->support = support;
    f->sing_order = sing_order;
    f->dn_ptr = (int *)malloc(sizeof(int) * (n + 1));
    memcpy(f->dn_ptr, dn_ptr, sizeof(int) * (n + 1));
    f->dn_idx = (int *)malloc(sizeof(int) * dn_ptr[n]);
    memcpy(f->dn_idx, dn_idx, sizeof(int) * dn_ptr[n]);
    f->X = (double *)malloc(sizeof(double) * 3 * n);
    memcpy(f->X, support, sizeof(double) * 3 * n);
    /* FEValues at the regular rule (computational_domain.cc:380-395) */
    const int nq = f->nq;
    double *pts = (double *)malloc(sizeof(double) * 2 * nq), *w = (double *)malloc(sizeof(double) * nq);
    orc_qgauss_2d(q_order, pts, w);
    f->qx = (double *)malloc(sizeof(double) * 3 * (size_t)nc * nq);
    f->qn = (double *)malloc(sizeof(double) * 3 * (size_t)nc * nq);
    f->qj = (double *)malloc(sizeof(double) * (size_t)nc * nq);
    f->qs = (double *)malloc(sizeof(double) * 4 * (size_t)nc * nq);
    for (int c = 0; c < nc; ++c) {
        double v12[12];
        for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) v12[3 * k + d] = support[3 * cell_dofs[4 * c + k] + d];
        const size_t o = (size_t)c * nq;
        orc_fe_values(v12, orient[c], nq, pts, w, f->qx + 3 * o, f->qn + 3 * o, f->qj + o,
                      f->qs + 4 * o);
    }
    free(pts);
    free(w);
    /* dof_to_elems */
    f->cell_ptr = (int *)calloc(n + 1, sizeof(int));
    for (int k = 0; k < 4 * nc; ++k) f->cell_ptr[cell_dofs[k] + 1]++;
    for (int i = 0; i < n; ++i) f->cell_ptr[i + 1] += f->cell_ptr[i];
    f->cell_items = (int *)malloc(sizeof(int) * 4 * nc);
    int *fill = (int *)malloc(sizeof(int) * n);
    memcpy(fill, f->cell_ptr, sizeof(int) * n);
    for (int c = 0; c < nc; ++c)
        for (int k = 0; k < 4; ++k) f->cell_items[fill[cell_dofs[4 * c + k]]++] = c;
    free(fill);
    build_octree(f);
    direct_integrals(f);
    multipole_integrals(f);
    /* block expansions */
    const int K = ncoef(f->p);
    f->mem = (double complex *)calloc((size_t)f->nb * 4 * K, sizeof(double complex));
    f->M1 = (Expansion *)calloc(f->nb, sizeof(Expansion));
    f->M2 = (Expansion *)calloc(f->nb, sizeof(Expansion));
    f->L1 = (Expansion *)calloc(f->nb, sizeof(Expansion));
    f->L2 = (Expansion *)calloc(f->nb, sizeof(Expansion));
    for (int b = 0; b < f->nb; ++b) {
        Expansion *E[4] = {&f->M1[b], &f->M2[b], &f->L1[b], &f->L2[b]};
        for (int t = 0; t < 4; ++t) {
            E[t]->c = f->mem + ((size_t)b * 4 + t) * K;
            for (int d = 0; d < 3; ++d) E[t]->ctr[d] = f->B[b].pmin[d] + f->B[b].delta / 2.;
        }
    }
    return f;
}

void orc_fma_destroy(orc_fma *f)
{
    if (!f) return;
    for (int b = 0; b < f->nb; ++b) {
        Block *blk = &f->B[b];
        iv_free(&blk->nodes);
        iv_free(&blk->qp);
        for (int s = 0; s < blk->nsub && blk->nn; ++s) {
            iv_free(&blk->nn[s]);
            iv_free(&blk->il[s]);
            iv_free(&blk->nil[s]);
        }
        free(blk->nn);
        free(blk->il);
        free(blk->nil);
    }
    free(f->B);
    for (int l = 0; l <= f->L; ++l) iv_free(&f->filled[l]);
    free(f->filled);
    iv_free(&f->childless);
    for (int i = 0; i < f->n && f->rows; ++i) {
        free(f->rows[i].col);
        free(f->rows[i].h);
        free(f->rows[i].g);
    }
    free(f->rows);
    free(f->start);
    free(f->end);
    free(f->qp2b);
    free(f->dn_ptr);
    free(f->dn_idx);
    free(f->X);
    free(f->qx);
    free(f->qn);
    free(f->qj);
    free(f->qs);
    free(f->cell_ptr);
    free(f->cell_items);
    free(f->el_leaf_ptr);
    free(f->el_cell);
    free(f->el_c);
    free(f->M1);
    free(f->M2);
    free(f->L1);
    free(f->L2);
    free(f->mem);
    free(f);
}

int orc_fma_products(orc_fma *f, const double *phi, const double *dphi_dn, double *prodN,
                     double *prodD)
{
    const int n = f->n, p = f->p, K = ncoef(p), L = f->L;
    memset(f->mem, 0, sizeof(double complex) * (size_t)f->nb * 4 * K);
    for (int b = 0; b < f->nb; ++b) f->M1[b].zero = f->M2[b].zero = f->L1[b].zero = f->L2[b].zero = 1;
    /* generate_multipole_expansions (bem_fma.cc:767-884) */
    for (int k = 0; k < f->childless.n; ++k) {
        const int b = f->childless.a[k];
        for (int e = f->el_leaf_ptr[b]; e < f->el_leaf_ptr[b + 1]; ++e) {
            const int cell = f->el_cell[e];
            for (int j = 0; j < 4; ++j) {
                const int dof = f->cd[4 * cell + j];
                const double complex *c1 = f->el_c + ((size_t)e * 8 + 2 * j) * K;
                const double complex *c2 = f->el_c + ((size_t)e * 8 + 2 * j + 1) * K;
                Expansion e2 = {(double complex *)c2, {0, 0, 0}, el_zero(c2, K)};
                Expansion e1 = {(double complex *)c1, {0, 0, 0}, el_zero(c1, K)};
                mp_add_scaled(&f->M2[b], p, &e2, dphi_dn[dof]);
                mp_add_scaled(&f->M1[b], p, &e1, phi[dof]);
            }
        }
    }
    for (int level = L; level > 0; --level)
        for (int kk = f->start[level]; kk <= f->end[level]; ++kk) {
            const int par = f->B[kk].parent;
            mp_translate(&f->M1[par], p, &f->M1[kk]);
            mp_translate(&f->M2[par], p, &f->M2[kk]);
        }
    /* multipole_matr_vect_products (bem_fma.cc:886-1087) */
    for (int i = 0; i < n; ++i) {
        double sh = 0.0, sg = 0.0;
        const SpRow *r = &f->rows[i];
        for (int k = 0; k < r->n; ++k) {
            sh += r->h[k] * phi[r->col[k]];
            sg += r->g[k] * dphi_dn[r->col[k]];
        }
        prodN[i] = sh;
        prodD[i] = sg;
    }
    for (int level = 1; level <= L; ++level) {
        const int sb = f->start[level], eb = f->end[level];
        for (int k = 0; k < f->filled[level].n; ++k) {
            const int jj = f->filled[level].a[k];
            const Block *b1 = &f->B[jj];
            lc_translate(&f->L1[jj], p, &f->L1[b1->parent]);
            lc_translate(&f->L2[jj], p, &f->L2[b1->parent]);
            for (int s = 0; s < b1->nsub && b1->nil; ++s) {
                const ivec *nil = &b1->nil[s];
                const int lo = lower_bound(nil, sb), hi = lower_bound(nil, eb + 1);
                for (int t = lo; t < hi; ++t) {
                    lc_add_mp(&f->L1[jj], p, &f->M1[nil->a[t]]);
                    lc_add_mp(&f->L2[jj], p, &f->M2[nil->a[t]]);
                }
                for (int t = hi; t < nil->n; ++t)
                    for (int i = 0; i < b1->nodes.n; ++i) {
                        const int nd = b1->nodes.a[i];
                        prodD[nd] += mp_eval(&f->M2[nil->a[t]], p, f->X + 3 * nd);
                        prodN[nd] += mp_eval(&f->M1[nil->a[t]], p, f->X + 3 * nd);
                    }
            }
        }
    }
    for (int k = 0; k < f->childless.n; ++k) {
        const int b = f->childless.a[k];
        const Block *b1 = &f->B[b];
        for (int i = 0; i < b1->nodes.n; ++i) {
            const int nd = b1->nodes.a[i];
            prodD[nd] += lc_eval(&f->L2[b], p, f->X + 3 * nd);
            prodN[nd] += lc_eval(&f->L1[b], p, f->X + 3 * nd);
        }
    }
    return 0;
}

/* How often each (collocation node, regular quadrature point) pair enters the product:
 * near field (a singular cell covers all its points once), M2L -> L2L -> L2P, or M2P.
 * count is n x (nc * nq), zeroed by the caller; every entry must end at 1. */
int orc_fma_coverage(const orc_fma *f, unsigned char *count)
{
    const int nq = f->nq, L = f->L;
    const size_t nqp = (size_t)f->nc * nq;
    int *buf = (int *)malloc(sizeof(int) * nqp);
    int *bl = (int *)malloc(sizeof(int) * (size_t)f->nb);
#define HIT(node, q) (count[(size_t)(node) * nqp + (q)]++)
    for (int k = 0; k < f->childless.n; ++k) {
        const Block *b1 = &f->B[f->childless.a[k]];
        if (b1->nodes.n == 0) continue;
        const ivec *il = &b1->il[b1->nsub - 1];
        const int m = collect_points(f, il->a, il->n, buf);
        for (int i = 0; i < b1->nodes.n; ++i) {
            const int node = b1->nodes.a[i];
            for (int s = 0; s < m;) {
                const int cell = buf[s] / nq;
                int e = s;
                while (e < m && buf[e] / nq == cell) ++e;
                int sing = 0;
                for (int j = 0; j < 4; ++j)
                    for (int p = f->dn_ptr[node]; p < f->dn_ptr[node + 1]; ++p)
                        sing |= f->dn_idx[p] == f->cd[4 * cell + j];
                if (sing)
                    for (int q = 0; q < nq; ++q) HIT(node, (size_t)cell * nq + q);
                else
                    for (int t = s; t < e; ++t) HIT(node, buf[t]);
                s = e;
            }
        }
    }
    for (int level = 1; level <= L; ++level) {
        const int sb = f->start[level], eb = f->end[level];
        for (int k = 0; k < f->filled[level].n; ++k) {
            const Block *b1 = &f->B[f->filled[level].a[k]];
            int nbl = 0;
            for (int s = 0; s < b1->nsub && b1->nil; ++s) {
                const ivec *nil = &b1->nil[s];
                const int lo = lower_bound(nil, sb);
                for (int t = 0; t < lo; ++t) bl[nbl++] = nil->a[t];
                /* M2L (same level) and M2P (finer levels): all points of the source block */
                for (int t = lo; t < nil->n; ++t) {
                    const ivec *qp = &f->B[nil->a[t]].qp;
                    for (int i = 0; i < b1->nodes.n; ++i)
                        for (int u = 0; u < qp->n; ++u) HIT(b1->nodes.a[i], qp->a[u]);
                }
                (void)eb;
            }
            if (nbl == 0) continue;
            const int m = collect_points(f, bl, nbl, buf);
            for (int i = 0; i < b1->nodes.n; ++i)
                for (int t = 0; t < m; ++t) HIT(b1->nodes.a[i], buf[t]);
        }
    }
#undef HIT
    free(buf);
    free(bl);
    return 0;
}

void orc_fma_info(const orc_fma *f, long *info)
{
    long nnz = 0, maxlev = 0;
    for (int i = 0; i < f->n; ++i) nnz += f->rows[i].n;
    for (int b = 0; b < f->nb; ++b)
        if (f->B[b].level > maxlev) maxlev = f->B[b].level;
    info[0] = f->nb;
    info[1] = f->childless.n;
    info[2] = nnz;
    info[3] = f->el_leaf_ptr[f->nb];
    info[4] = maxlev;
}

/* Expansion primitives for closed-form checks.  kind: 0 point charge (strength 1 at src)
 * about c[0] evaluated at x by M2P; 1 the normal dipole (normal at src) likewise; 2 the
 * charge expanded about c[0], translated to c[1] (M2M), evaluated at x; 3 the charge about
 * c[0], converted to a local expansion about c[1] (M2L), evaluated at x (L2P); 4 as 3,
 * then translated to c[2] (L2L) before L2P. */
int orc_fma_expansion_check(int p, int kind, const double *src, const double *normal,
                            const double *c, const double *x, double *out)
{
    if (p < 0 || p > FMA_PMAX) return -1;
    init_tables();
    const int K = ncoef(p);
    double complex *mem = (double complex *)calloc(3 * K, sizeof(double complex));
    Expansion a = {mem, {c[0], c[1], c[2]}, 1};
    Expansion b = {mem + K, {c[3], c[4], c[5]}, 1};
    Expansion d = {mem + 2 * K, {c[6], c[7], c[8]}, 1};
    if (kind == 1)
        mp_add_normder(&a, p, 1.0, src, normal);
    else
        mp_add_point(&a, p, 1.0, src);
    switch (kind) {
    case 0:
    case 1:
        *out = mp_eval(&a, p, x);
        break;
    case 2:
        mp_translate(&b, p, &a);
        *out = mp_eval(&b, p, x);
        break;
    case 3:
        lc_add_mp(&b, p, &a);
        *out = lc_eval(&b, p, x);
        break;
    case 4:
        lc_add_mp(&b, p, &a);
        lc_translate(&d, p, &b);
        *out = lc_eval(&d, p, x);
        break;
    default:
        free(mem);
        return -1;
    }
    free(mem);
    return 0;
}

/* normalised associated Legendre function and its sin-weighted derivative (tests) */
double orc_fma_legendre(int n, int m, double x, double *dsin)
{
    if (n < 0 || n > FMA_LMAX || m < 0 || m > n) return 0.0;
    init_tables();
    double P[LS * LS], dS[LS * LS];
    legtab(n, x, P, dS);
    if (dsin) *dsin = dS[n * LS + m];
    return P[n * LS + m];
}
