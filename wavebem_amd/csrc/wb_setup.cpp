// wb_setup.cpp -- host-side preparation for the device BEM path.
//
//  * quadrature tables with the semantics of deal.II QGauss<2> and
//    QGaussOneOverR<2>(n, vertex, true) as requested by BEMProblem::assemble_system
//    (source/bem_problem.cc:116-137);
//  * mesh planning: double-node classes (source/computational_domain.cc:257-307),
//    the device dof ordering, column tiles and the singular (row, cell) pairs of
//    source/bem_problem.cc:214-230;
//  * the constraint rules of BEMProblem::compute_constraints
//    (source/bem_problem.cc:990-1105) closed like deal.II's ConstraintMatrix::close.
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <unordered_map>

#include "wb_internal.h"

namespace wb {

// Gauss-Legendre nodes on [0,1], ascending: Newton's method on the three-term Legendre
// recurrence in long double, started from the Tricomi-type estimate
// cos(pi (i - 1/4)/(n + 1/2)); weights 1/((1 - z^2) P_n'(z)^2) scaled to [0,1].  This is
// the algorithm deal.II uses for QGauss<1>, so nodes agree to the last bit or two.
void gauss_legendre_01(int n, double *x, double *w)
{
    const int m = (n + 1) / 2;
    const long double tol = std::max((long double)__DBL_EPSILON__ / 100,
                                     (long double)__LDBL_EPSILON__ * 5);
    for (int i = 1; i <= m; ++i) {
        long double z = std::cos((long double)M_PI * (i - 0.25L) / (n + 0.5L));
        long double dp = 0, pn = 0;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, pm = 0.0L;
            for (int j = 0; j < n; ++j) {
                const long double pj = ((2.0L * j + 1.0L) * z * p0 - j * pm) / (j + 1);
                pm = p0;
                p0 = pj;
            }
            pn = p0;
            dp = n * (z * pn - pm) / (z * z - 1.0L);
            const long double dz = pn / dp;
            z -= dz;
            if (std::fabs(dz) <= tol) break;
        }
        const double half = (double)(0.5L * z);
        const double wt = (double)(1.0L / ((1.0L - z * z) * dp * dp));
        x[i - 1] = 0.5 - half;
        x[n - i] = 0.5 + half;
        w[i - 1] = wt;
        w[n - i] = wt;
    }
}

void build_quadrature(int q_order, int s_order, Quadrature &Q)
{
    Q.q_order = q_order;
    Q.s_order = s_order;
    double gx[64], gw[64];
    gauss_legendre_01(q_order, gx, gw);
    const int nq = q_order * q_order;
    Q.reg_pts.resize(2 * nq);
    Q.reg_w.resize(nq);
    for (int j = 0; j < q_order; ++j)        // second coordinate slowest
        for (int i = 0; i < q_order; ++i) {  // first coordinate fastest
            const int q = i + q_order * j;
            Q.reg_pts[2 * q] = gx[i];
            Q.reg_pts[2 * q + 1] = gx[j];
            Q.reg_w[q] = gw[i] * gw[j];
        }
    // singular rule for vertex 0: polar-type map of QGauss<2>(s) to the triangle under
    // the diagonal, mirrored across it; weight carries |p| (1/R factored out).
    gauss_legendre_01(s_order, gx, gw);
    const int ng = s_order * s_order, ns = 2 * ng;
    std::vector<double> p0(2 * ns), w0(ns);
    const double pi4 = M_PI / 4;
    for (int j = 0; j < s_order; ++j)
        for (int i = 0; i < s_order; ++i) {
            const int q = i + s_order * j;
            const double u = gx[i], v = gx[j];
            const double a = u, b = u * std::tan(pi4 * v);
            double wq = gw[i] * gw[j] * pi4 / std::cos(pi4 * v);
            wq *= std::sqrt(a * a + b * b);
            p0[2 * q] = a;
            p0[2 * q + 1] = b;
            w0[q] = wq;
            p0[2 * (ng + q)] = b;
            p0[2 * (ng + q) + 1] = a;
            w0[ng + q] = wq;
        }
    Q.sing_pts.resize(4 * 2 * ns);
    Q.sing_w.resize(4 * ns);
    const double theta[4] = {0.0, M_PI / 2, -M_PI / 2, M_PI};
    for (int v = 0; v < 4; ++v) {
        const double c = std::cos(theta[v]), s = std::sin(theta[v]);
        for (int q = 0; q < ns; ++q) {
            double x = p0[2 * q], y = p0[2 * q + 1];
            if (v != 0) {  // rotate about (1/2, 1/2)
                const double xx = x - 0.5, yy = y - 0.5;
                x = c * xx - s * yy + 0.5;
                y = s * xx + c * yy + 0.5;
            }
            Q.sing_pts[2 * (v * ns + q)] = x;
            Q.sing_pts[2 * (v * ns + q) + 1] = y;
            Q.sing_w[v * ns + q] = w0[q];
        }
    }
}

static uint64_t spread3(uint64_t v)
{
    v &= 0x1fffffULL;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

struct UnionFind {
    std::vector<int> p;
    explicit UnionFind(int n) : p(n) { std::iota(p.begin(), p.end(), 0); }
    int find(int x)
    {
        while (p[x] != x) {
            p[x] = p[p[x]];
            x = p[x];
        }
        return x;
    }
    void unite(int a, int b)
    {
        a = find(a);
        b = find(b);
        if (a != b) p[std::max(a, b)] = std::min(a, b);
    }
};

// Geometric coincidence (distance < 1e-8) via a hash grid of cell size 1e-8: coincident
// points differ by at most one cell per axis.  Same sets as the reference's O(n^2) scan.
static void double_nodes(int n, const double *X, std::vector<int> &ptr, std::vector<int> &idx)
{
    const double tol = 1e-8;
    std::unordered_map<uint64_t, std::vector<int>> grid;
    grid.reserve(2 * n);
    auto key = [](int64_t a, int64_t b, int64_t c) {
        return (uint64_t)(a * 73856093LL) ^ (uint64_t)(b * 19349663LL) ^ (uint64_t)(c * 83492791LL);
    };
    std::vector<int64_t> q(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) q[3 * i + d] = (int64_t)std::floor(X[3 * i + d] / tol);
        grid[key(q[3 * i], q[3 * i + 1], q[3 * i + 2])].push_back(i);
    }
    std::vector<std::vector<int>> sets(n);
    for (int i = 0; i < n; ++i) {
        std::vector<int> &s = sets[i];
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    auto it = grid.find(key(q[3 * i] + dx, q[3 * i + 1] + dy, q[3 * i + 2] + dz));
                    if (it == grid.end()) continue;
                    for (int j : it->second) {
                        const double a = X[3 * i] - X[3 * j], b = X[3 * i + 1] - X[3 * j + 1],
                                     c = X[3 * i + 2] - X[3 * j + 2];
                        if (std::sqrt(a * a + b * b + c * c) < tol) s.push_back(j);
                    }
                }
        std::sort(s.begin(), s.end());
        s.erase(std::unique(s.begin(), s.end()), s.end());
    }
    ptr.assign(n + 1, 0);
    idx.clear();
    for (int i = 0; i < n; ++i) {
        idx.insert(idx.end(), sets[i].begin(), sets[i].end());
        ptr[i + 1] = (int)idx.size();
    }
}

// ------------------------------------------------------------------- strip planning
namespace {

// Integer grid coordinates (u, v) of the dofs of one mesh component whose cells form a
// logically structured quad grid: no cell repeats a dof, an edge belongs to at most two
// cells, the coordinates carried across shared edges agree at every vertex, every cell is
// reached through edges and is a unit square.  Returns false otherwise (irregular
// vertices, a grid closing on itself, cells meeting only at a corner).
bool grid_coords(const std::vector<int> &cells, const int32_t *cd, std::vector<int> &gu,
                 std::vector<int> &gv, std::vector<int> &mark, int stamp)
{
    static const int cyc[4] = {0, 1, 3, 2};  // cyclic vertex order of a deal.II quad
    std::unordered_map<uint64_t, std::pair<int, int>> edges;
    edges.reserve(4 * cells.size());
    auto ekey = [](int a, int b) {
        if (a > b) std::swap(a, b);
        return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
    };
    for (int c : cells) {
        const int32_t *v = cd + 4 * (size_t)c;
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j)
                if (v[i] == v[j]) return false;
        for (int i = 0; i < 4; ++i) {
            auto it = edges.emplace(ekey(v[cyc[i]], v[cyc[(i + 1) & 3]]), std::make_pair(c, -1));
            if (!it.second) {
                if (it.first->second.second >= 0) return false;
                it.first->second.second = c;
            }
        }
    }
    auto set = [&](int d, int u, int w) {
        if (gu[d] == INT_MIN) {
            gu[d] = u;
            gv[d] = w;
            return true;
        }
        return gu[d] == u && gv[d] == w;
    };
    const int c0 = cells[0];
    {
        const int32_t *v = cd + 4 * (size_t)c0;
        set(v[0], 0, 0);
        set(v[1], 1, 0);
        set(v[2], 0, 1);
        set(v[3], 1, 1);
    }
    std::vector<int> queue{c0};
    mark[c0] = stamp;
    for (size_t qi = 0; qi < queue.size(); ++qi) {
        const int c = queue[qi];
        const int32_t *v = cd + 4 * (size_t)c;
        for (int i = 0; i < 4; ++i) {
            // edge (p, q) of c; s is the other neighbour of p in c, so s - p points into c
            const int p = v[cyc[i]], q = v[cyc[(i + 1) & 3]], s = v[cyc[(i + 3) & 3]];
            const auto &e = edges[ekey(p, q)];
            const int o = e.first == c ? e.second : e.first;
            if (o < 0 || mark[o] == stamp) continue;
            const int eu = gu[s] - gu[p], ev = gv[s] - gv[p];
            const int32_t *x = cd + 4 * (size_t)o;
            int j = 0;
            while (j < 4 && x[cyc[j]] != p) ++j;
            if (j == 4) return false;
            int pa, qa;  // neighbours of p and q in o across the edge
            if (x[cyc[(j + 1) & 3]] == q) {
                pa = x[cyc[(j + 3) & 3]];
                qa = x[cyc[(j + 2) & 3]];
            } else if (x[cyc[(j + 3) & 3]] == q) {
                pa = x[cyc[(j + 1) & 3]];
                qa = x[cyc[(j + 2) & 3]];
            } else {
                return false;
            }
            if (!set(pa, gu[p] - eu, gv[p] - ev) || !set(qa, gu[q] - eu, gv[q] - ev)) return false;
            mark[o] = stamp;
            queue.push_back(o);
        }
    }
    if (queue.size() != cells.size()) return false;
    for (int c : cells) {
        const int32_t *v = cd + 4 * (size_t)c;
        const int au = gu[v[1]] - gu[v[0]], av = gv[v[1]] - gv[v[0]];
        const int bu = gu[v[2]] - gu[v[0]], bv = gv[v[2]] - gv[v[0]];
        if (std::abs(au) + std::abs(av) != 1 || std::abs(bu) + std::abs(bv) != 1 ||
            au * bu + av * bv != 0)
            return false;
        if (gu[v[3]] != gu[v[1]] + bu || gv[v[3]] != gv[v[1]] + bv) return false;
    }
    return true;
}

struct StripBuild {
    std::vector<int> order;                  // the component's dofs in device order
    std::vector<std::array<int, 3>> tasks;   // {first gap, gaps, cells}
    std::vector<std::array<int, 3>> gaps;    // {cell mask, first column (index into order), line mask}
    std::vector<std::array<int, STRIP_W + 1>> gcells;  // cells of a gap by position, -1: none
    std::vector<std::array<int, 5>> canon;   // {cell, canonical vertices 0..3 (dofs)}
    std::vector<int> strip_first;            // index into order where each strip starts
    long cells = 0;                          // cell evaluations of the schedule
};

// Strips of one structured component, positions along u (o = 0) or v (o = 1), tasks of
// about `target` entries.  False if two dofs or two cells share a grid point.
bool build_strips(const std::vector<int> &cells, const std::vector<int> &dofs,
                  const int32_t *cd, const std::vector<int> &gu, const std::vector<int> &gv,
                  int o, int target, StripBuild &B)
{
    B = StripBuild();
    auto P = [&](int d) { return o ? gv[d] : gu[d]; };
    auto A = [&](int d) { return o ? gu[d] : gv[d]; };
    int pmin = INT_MAX, pmax = INT_MIN, amin = INT_MAX, amax = INT_MIN;
    for (int d : dofs) {
        pmin = std::min(pmin, P(d));
        pmax = std::max(pmax, P(d));
        amin = std::min(amin, A(d));
        amax = std::max(amax, A(d));
    }
    const int Np = pmax - pmin + 1, Na = amax - amin + 1;
    if ((double)Np * Na > 64.0 * (double)dofs.size() + 1024) return false;  // too sparse a grid
    std::vector<int> node((size_t)Np * Na, -1), cellg((size_t)Np * Na, -1);
    for (int d : dofs) {
        int &slot = node[(size_t)(P(d) - pmin) + (size_t)Np * (A(d) - amin)];
        if (slot >= 0) return false;
        slot = d;
    }
    for (int c : cells) {
        int p0 = INT_MAX, a0 = INT_MAX;
        for (int k = 0; k < 4; ++k) {
            p0 = std::min(p0, P(cd[4 * (size_t)c + k]) - pmin);
            a0 = std::min(a0, A(cd[4 * (size_t)c + k]) - amin);
        }
        int &slot = cellg[(size_t)p0 + (size_t)Np * a0];
        if (slot >= 0) return false;
        slot = c;
        std::array<int, 5> cn{c, node[(size_t)p0 + (size_t)Np * a0],
                              node[(size_t)p0 + 1 + (size_t)Np * a0],
                              node[(size_t)p0 + (size_t)Np * (a0 + 1)],
                              node[(size_t)p0 + 1 + (size_t)Np * (a0 + 1)]};
        B.canon.push_back(cn);
    }
    const int ns = (Np + STRIP_W - 1) / STRIP_W;
    for (int s = 0; s < ns; ++s) {
        B.strip_first.push_back((int)B.order.size());
        const int b0 = (int)((long)s * Np / ns), w = (int)((long)(s + 1) * Np / ns) - b0;
        auto cell_at = [&](int g, int k) {
            const int p0 = b0 + k - 1;
            if (p0 < 0 || p0 > Np - 2 || g < 0 || g > Na - 2) return -1;
            return cellg[(size_t)p0 + (size_t)Np * g];
        };
        long total = 0;
        for (int g = 0; g + 1 < Na; ++g)
            for (int k = 0; k <= w; ++k) total += cell_at(g, k) >= 0;
        int np = std::max(1, (int)std::lround((double)total / std::max(1, target)));
        np = std::min(np, Na);
        for (int r = 0; r < np; ++r) {
            // levels [a0, a1): gap t (task-local, -1 .. L-1) holds the cells between lines
            // t and t + 1; after it, line t is complete
            const int a0 = (int)((long)r * Na / np), a1 = (int)((long)(r + 1) * Na / np);
            const int L = a1 - a0;
            const long cells_before = B.cells;
            B.tasks.push_back({(int)B.gaps.size(), L + 1, 0});
            for (int t = -1; t < L; ++t) {
                std::array<int, STRIP_W + 1> gc;
                int cmask = 0;
                for (int k = 0; k <= STRIP_W; ++k) {
                    gc[k] = k <= w ? cell_at(a0 + t, k) : -1;
                    if (gc[k] >= 0) {
                        cmask |= 1 << k;
                        ++B.cells;
                    }
                }
                int first = -1, lmask = 0;
                if (t >= 0)
                    for (int p = 0; p < w; ++p) {
                        const int d = node[(size_t)(b0 + p) + (size_t)Np * (a0 + t)];
                        if (d < 0) continue;
                        if (first < 0) first = (int)B.order.size();
                        lmask |= 1 << p;
                        B.order.push_back(d);
                    }
                B.gaps.push_back({cmask, first < 0 ? 0 : first, lmask});
                B.gcells.push_back(gc);
            }
            B.tasks.back()[2] = (int)(B.cells - cells_before);
        }
    }
    return true;
}

}  // namespace

std::string plan_mesh(int n, int nc, const double *X, const int32_t *cd, const int32_t *cmat,
                      const int32_t *dnp, const int32_t *dni, int tc, int strips, MeshPlan &P)
{
    P = MeshPlan();
    if (tc < 1 || tc > TILE_COLS) return "tile width must be 1..16";
    P.tc = tc;
    P.n = n;
    P.nc = nc;
    for (int c = 0; c < nc; ++c)
        for (int k = 0; k < 4; ++k)
            if (cd[4 * c + k] < 0 || cd[4 * c + k] >= n) return "cell dof index out of range";
    for (int i = 0; i < 3 * n; ++i)
        if (!std::isfinite(X[i])) return "non-finite support point";
    if (dnp && dni) {
        P.dn_ptr.assign(dnp, dnp + n + 1);
        if (P.dn_ptr[0] != 0) return "double node CSR must start at 0";
        for (int i = 0; i < n; ++i)
            if (P.dn_ptr[i + 1] < P.dn_ptr[i]) return "double node CSR pointer decreasing";
        P.dn_idx.assign(dni, dni + P.dn_ptr[n]);
        for (int v : P.dn_idx)
            if (v < 0 || v >= n) return "double node index out of range";
    } else {
        double_nodes(n, X, P.dn_ptr, P.dn_idx);
    }
    // classes = connected components; every set must be exactly its class
    UnionFind uf(n);
    for (int i = 0; i < n; ++i)
        for (int p = P.dn_ptr[i]; p < P.dn_ptr[i + 1]; ++p) uf.unite(i, P.dn_idx[p]);
    std::vector<int> csize(n, 0);
    for (int i = 0; i < n; ++i) csize[uf.find(i)]++;
    for (int i = 0; i < n; ++i) {
        bool self = false;
        for (int p = P.dn_ptr[i]; p < P.dn_ptr[i + 1]; ++p) self |= P.dn_idx[p] == i;
        if (!self) return "double_nodes_set[i] must contain i";
        if (P.dn_ptr[i + 1] - P.dn_ptr[i] != csize[uf.find(i)])
            return "double node sets are not an equivalence relation";
    }
    // dof material: first cell containing the dof
    std::vector<int> dmat(n, 0);
    if (cmat) {
        std::vector<char> seen(n, 0);
        for (int c = 0; c < nc; ++c)
            for (int k = 0; k < 4; ++k) {
                const int d = cd[4 * c + k];
                if (!seen[d]) {
                    seen[d] = 1;
                    dmat[d] = cmat[c];
                }
            }
    }
    double lo[3] = {X[0], X[1], X[2]}, hi[3] = {X[0], X[1], X[2]};
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], X[3 * i + d]);
            hi[d] = std::max(hi[d], X[3 * i + d]);
        }
    double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    if (ext <= 0) ext = 1;
    std::vector<uint64_t> key(n);
    for (int i = 0; i < n; ++i) {
        uint64_t m = 0;
        for (int d = 0; d < 3; ++d) {
            double t = (X[3 * i + d] - lo[d]) / ext;
            t = std::min(std::max(t, 0.0), 1.0);
            m |= spread3((uint64_t)(t * 2097151.0)) << d;
        }
        key[i] = m;
    }
    std::vector<int> seeds(n);
    std::iota(seeds.begin(), seeds.end(), 0);
    std::stable_sort(seeds.begin(), seeds.end(), [&](int a, int b) {
        if (dmat[a] != dmat[b]) return dmat[a] < dmat[b];
        return key[a] < key[b];
    });
    // Column tiles by greedy region growing on the mesh graph: a tile starts next to the
    // previous one (or at the first free dof in material/Morton order) and repeatedly
    // takes the neighbouring dof that brings the fewest new cells into the tile.  Cells
    // touching a tile are evaluated once per tile, so compact tiles mean less redundant
    // kernel work.  The device order is the concatenation of the tiles.
    std::vector<int> dc_ptr(n + 1, 0), dc_idx(4 * (size_t)nc);
    for (int c = 0; c < nc; ++c)
        for (int k = 0; k < 4; ++k) dc_ptr[cd[4 * c + k] + 1]++;
    for (int i = 0; i < n; ++i) dc_ptr[i + 1] += dc_ptr[i];
    {
        std::vector<int> fill(dc_ptr.begin(), dc_ptr.end() - 1);
        for (int c = 0; c < nc; ++c)
            for (int k = 0; k < 4; ++k) dc_idx[fill[cd[4 * c + k]]++] = c;
    }
    std::vector<char> assigned(n, 0);
    std::vector<int> stamp(nc, -1);
    std::vector<int> order;
    order.reserve(n);
    // Structured patches first: each connected component of the cell graph whose cells
    // form a logically structured grid is laid out as strips (STRIP_W-wide bands of grid
    // positions, level by level), which the strip kernel sweeps with every cell evaluated
    // once per strip it touches.  Components are taken in material / Morton order.
    P.cell_int_s.assign(cd, cd + 4 * (size_t)nc);  // caller numbering until perm is known
    P.canon_sign.assign(nc, 1);
    std::vector<std::array<int, 3>> all_tasks;
    if (strips) {
        UnionFind cu(n);
        for (int c = 0; c < nc; ++c)
            for (int k = 1; k < 4; ++k) cu.unite(cd[4 * c], cd[4 * c + k]);
        std::vector<int> root_rank(n, -1), comp_of(n);
        int n_comp = 0;
        for (int s : seeds) {
            const int r = cu.find(s);
            if (root_rank[r] < 0) root_rank[r] = n_comp++;
        }
        std::vector<std::vector<int>> ccells(n_comp), cdofs(n_comp);
        for (int i = 0; i < n; ++i) cdofs[root_rank[cu.find(i)]].push_back(i);
        for (int c = 0; c < nc; ++c) ccells[root_rank[cu.find(cd[4 * c])]].push_back(c);
        std::vector<int> gu(n, INT_MIN), gv(n, INT_MIN), mark(nc, -1);
        StripBuild B0, B1;
        for (int k = 0; k < n_comp; ++k) {
            if (ccells[k].empty()) continue;
            if (!grid_coords(ccells[k], cd, gu, gv, mark, k)) continue;
            const bool ok0 = build_strips(ccells[k], cdofs[k], cd, gu, gv, 0, STRIP_TASK_ENTRIES, B0);
            const bool ok1 = build_strips(ccells[k], cdofs[k], cd, gu, gv, 1, STRIP_TASK_ENTRIES, B1);
            if (!ok0 && !ok1) continue;
            const StripBuild &B = (ok0 && (!ok1 || B0.cells <= B1.cells)) ? B0 : B1;
            const int base = (int)order.size();
            const int goff = (int)P.st_gap.size() / 4;
            for (int f : B.strip_first) P.run_start.push_back(base + f);
            for (int d : B.order) {
                order.push_back(d);
                assigned[d] = 1;
            }
            for (const auto &g : B.gaps) {
                P.st_gap.insert(P.st_gap.end(), {g[0], g[2] ? base + g[1] : 0, g[2], 0});
            }
            for (const auto &gc : B.gcells) P.st_gcell.insert(P.st_gcell.end(), gc.begin(), gc.end());
            for (const auto &t : B.tasks) all_tasks.push_back({t[0] + goff, t[1], t[2]});
            for (const auto &cn : B.canon) {
                const int c = cn[0];
                int kk[4];
                for (int i = 0; i < 4; ++i) {
                    P.cell_int_s[4 * (size_t)c + i] = cn[1 + i];
                    kk[i] = 0;
                    while (cd[4 * (size_t)c + kk[i]] != cn[1 + i]) ++kk[i];
                }
                // orientation of the canonical reference frame in the cell's own
                const int e1x = (kk[1] & 1) - (kk[0] & 1), e1y = (kk[1] >> 1) - (kk[0] >> 1);
                const int e2x = (kk[2] & 1) - (kk[0] & 1), e2y = (kk[2] >> 1) - (kk[0] >> 1);
                P.canon_sign[c] = (int8_t)(e1x * e2y - e1y * e2x);
            }
        }
    }
    P.strip_dofs = (int)order.size();
    // longest tasks first, so that the short ones fill the end of the grid
    std::stable_sort(all_tasks.begin(), all_tasks.end(), [](const std::array<int, 3> &a,
                                                            const std::array<int, 3> &b) {
        return a[2] > b[2];
    });
    P.n_tasks = (int)all_tasks.size();
    P.strip_cells = 0;
    for (const auto &t : all_tasks) {
        P.st_task.insert(P.st_task.end(), {t[0], t[1]});
        P.strip_cells += t[2];
    }
    std::vector<int> tstart{(int)order.size()};
    std::vector<int> prev, tile, front;
    size_t seed_ptr = 0;
    auto free_degree = [&](int v) {
        int d = 0;
        for (int p = dc_ptr[v]; p < dc_ptr[v + 1]; ++p)
            for (int k = 0; k < 4; ++k) d += !assigned[cd[4 * dc_idx[p] + k]];
        return d;
    };
    auto push_front = [&](int v) {
        for (int p = dc_ptr[v]; p < dc_ptr[v + 1]; ++p)
            for (int k = 0; k < 4; ++k) {
                const int w = cd[4 * dc_idx[p] + k];
                if (!assigned[w] && std::find(front.begin(), front.end(), w) == front.end())
                    front.push_back(w);
            }
    };
    while ((int)order.size() < n) {
        const int tid = (int)tstart.size() - 1;
        int seed = -1;
        long best = 0;
        for (int t : prev)
            for (int p = dc_ptr[t]; p < dc_ptr[t + 1]; ++p)
                for (int k = 0; k < 4; ++k) {
                    const int w = cd[4 * dc_idx[p] + k];
                    if (assigned[w]) continue;
                    const long s = (long)free_degree(w);
                    if (seed < 0 || s < best || (s == best && key[w] < key[seed])) {
                        seed = w;
                        best = s;
                    }
                }
        if (seed < 0) {
            while (assigned[seeds[seed_ptr]]) ++seed_ptr;
            int looked = 0;
            for (size_t q = seed_ptr; q < seeds.size() && looked < 32; ++q) {
                const int w = seeds[q];
                if (assigned[w]) continue;
                if (dmat[w] != dmat[seeds[seed_ptr]]) break;
                ++looked;
                const long s = (long)free_degree(w);
                if (seed < 0 || s < best) {
                    seed = w;
                    best = s;
                }
            }
        }
        tile.assign(1, seed);
        assigned[seed] = 1;
        for (int p = dc_ptr[seed]; p < dc_ptr[seed + 1]; ++p) stamp[dc_idx[p]] = tid;
        front.clear();
        push_front(seed);
        while ((int)tile.size() < tc && !front.empty()) {
            int bi = -1, bnew = 0, bin = 0;
            for (int f = 0; f < (int)front.size(); ++f) {
                const int v = front[f];
                int nw = 0, in = 0;
                for (int p = dc_ptr[v]; p < dc_ptr[v + 1]; ++p) {
                    const int c = dc_idx[p];
                    if (stamp[c] != tid) ++nw;
                    for (int k = 0; k < 4; ++k) {
                        const int w = cd[4 * c + k];
                        in += (w != v && assigned[w] &&
                               std::find(tile.begin(), tile.end(), w) != tile.end());
                    }
                }
                if (bi < 0 || nw < bnew || (nw == bnew && in > bin) ||
                    (nw == bnew && in == bin && key[v] < key[front[bi]])) {
                    bi = f;
                    bnew = nw;
                    bin = in;
                }
            }
            const int v = front[bi];
            front.erase(front.begin() + bi);
            tile.push_back(v);
            assigned[v] = 1;
            for (int p = dc_ptr[v]; p < dc_ptr[v + 1]; ++p) stamp[dc_idx[p]] = tid;
            push_front(v);
            front.erase(std::remove_if(front.begin(), front.end(), [&](int w) { return assigned[w]; }),
                        front.end());
        }
        order.insert(order.end(), tile.begin(), tile.end());
        tstart.push_back((int)order.size());
        prev = tile;
    }
    P.iperm = order;
    P.perm.assign(n, 0);
    for (int i = 0; i < n; ++i) P.perm[order[i]] = i;
    P.ld = ((size_t)n + 63) / 64 * 64;
    P.cls.resize(n);
    for (int i = 0; i < n; ++i) P.cls[i] = P.perm[uf.find(P.iperm[i])];
    P.cell_int.resize(4 * (size_t)nc);
    P.cell_cls.resize(4 * (size_t)nc);
    for (int c = 0; c < nc; ++c)
        for (int k = 0; k < 4; ++k) {
            const int d = cd[4 * c + k];
            P.cell_int[4 * c + k] = P.perm[d];
            P.cell_cls[4 * c + k] = P.perm[uf.find(d)];
            P.cell_int_s[4 * c + k] = P.perm[P.cell_int_s[4 * c + k]];
        }
    // column tiles: [tile_start[t], tile_start[t+1]) in the device order, <= tc wide
    P.n_tiles = (int)tstart.size() - 1;
    P.tile_start = tstart;
    std::vector<int> tile_of(n, -1);  // -1: a strip column
    for (int t = 0; t < P.n_tiles; ++t)
        for (int i = tstart[t]; i < tstart[t + 1]; ++i) tile_of[i] = t;
    std::vector<std::vector<int>> tl(P.n_tiles), tpk(P.n_tiles);
    for (int c = 0; c < nc; ++c) {
        int tiles[4];
        for (int k = 0; k < 4; ++k) tiles[k] = tile_of[P.cell_int[4 * c + k]];
        for (int k = 0; k < 4; ++k) {
            bool dup = tiles[k] < 0;
            for (int j = 0; j < k; ++j) dup |= tiles[j] == tiles[k];
            if (dup) continue;
            const int t = tiles[k];
            uint32_t packed = 0;
            for (int j = 0; j < 4; ++j) {
                const int v = P.cell_int[4 * c + j];
                const int local = (tile_of[v] == t) ? v - tstart[t] : -1;
                packed |= (uint32_t)(uint8_t)(int8_t)local << (8 * j);
            }
            tl[t].push_back(c);
            tpk[t].push_back((int)packed);
        }
    }
    P.tile_ptr.assign(P.n_tiles + 1, 0);
    for (int t = 0; t < P.n_tiles; ++t) {
        P.tile_ptr[t + 1] = P.tile_ptr[t] + (int)tl[t].size();
        P.tile_cell.insert(P.tile_cell.end(), tl[t].begin(), tl[t].end());
        P.tile_cols.insert(P.tile_cols.end(), tpk[t].begin(), tpk[t].end());
    }
    // singular pairs: cells with a vertex in the row's class, first such vertex
    std::vector<std::vector<int>> class_cells(n);
    for (int c = 0; c < nc; ++c)
        for (int k = 0; k < 4; ++k) {
            auto &v = class_cells[P.cell_cls[4 * c + k]];
            if (v.empty() || v.back() != c) v.push_back(c);
        }
    P.sing_ptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        auto &v = class_cells[P.cls[i]];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        for (int c : v) {
            int jstar = -1;
            for (int k = 0; k < 4; ++k)
                if (P.cell_cls[4 * c + k] == P.cls[i]) {
                    jstar = k;
                    break;
                }
            P.sing_cell.push_back(c);
            P.sing_vtx.push_back(jstar);
        }
        P.sing_ptr[i + 1] = (int)P.sing_cell.size();
    }
    return "";
}

// ------------------------------------------------------------------- constraints
namespace {
struct Line {
    std::vector<std::pair<int, double>> e;
    double inhom = 0;
};
}  // namespace

// ------------------------------------------------------------------- compressed operator plan
void plan_hmat(const MeshPlan &P, const double *X, HmatPlan &H)
{
    H = HmatPlan();
    const int n = P.n;
    // segments: pieces of at most HM_SEG of each strip, then groups of whole column tiles
    std::vector<int> seg_a, seg_b;
    std::vector<int> starts = P.run_start;
    std::sort(starts.begin(), starts.end());
    starts.push_back(P.strip_dofs);
    for (size_t k = 0; k + 1 < starts.size(); ++k)
        for (int a = starts[k]; a < starts[k + 1]; a += HM_SEG) {
            seg_a.push_back(a);
            seg_b.push_back(std::min(a + HM_SEG, starts[k + 1]));
        }
    {
        int a = P.strip_dofs;
        for (int t = 0; t < P.n_tiles; ++t) {
            const int e = P.tile_start[t + 1];
            if (e - a > HM_SEG && P.tile_start[t] > a) {
                seg_a.push_back(a);
                seg_b.push_back(P.tile_start[t]);
                a = P.tile_start[t];
            }
        }
        if (a < n) {
            seg_a.push_back(a);
            seg_b.push_back(n);
        }
    }
    const int ns = (int)seg_a.size();
    std::vector<std::array<double, 3>> cen(ns);
    for (int k = 0; k < ns; ++k) {
        double c[3] = {0, 0, 0};
        for (int i = seg_a[k]; i < seg_b[k]; ++i)
            for (int d = 0; d < 3; ++d) c[d] += X[3 * (size_t)i + d];
        for (int d = 0; d < 3; ++d) cen[k][d] = c[d] / (seg_b[k] - seg_a[k]);
    }
    // recursive bisection of the segments along the longest extent of their centroids;
    // the left part takes a multiple of HM_SEGS segments, so that most leaves are full
    std::vector<int> ids(ns);
    for (int k = 0; k < ns; ++k) ids[k] = k;
    std::vector<std::vector<int>> leaves;
    std::vector<std::pair<int, int>> stack{{0, ns}};
    while (!stack.empty()) {
        auto [lo, hi] = stack.back();
        stack.pop_back();
        const int cnt = hi - lo;
        if (cnt <= HM_SEGS) {
            leaves.emplace_back(ids.begin() + lo, ids.begin() + hi);
            continue;
        }
        double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
        for (int k = lo; k < hi; ++k)
            for (int d = 0; d < 3; ++d) {
                mn[d] = std::min(mn[d], cen[ids[k]][d]);
                mx[d] = std::max(mx[d], cen[ids[k]][d]);
            }
        int ax = 0;
        for (int d = 1; d < 3; ++d)
            if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
        std::stable_sort(ids.begin() + lo, ids.begin() + hi,
                         [&](int a, int b) { return cen[a][ax] < cen[b][ax]; });
        const int left = HM_SEGS * std::max(1, (cnt + HM_SEGS) / (2 * HM_SEGS));
        stack.push_back({lo + left, hi});
        stack.push_back({lo, lo + left});
    }
    const int ncl = (int)leaves.size();
    H.ncl = ncl;
    H.cl_idx.assign((size_t)ncl * HM_CL, -1);
    H.cl_size.assign(ncl, 0);
    H.cl_seg.assign((size_t)ncl * HM_SEGS * 2, 0);
    std::vector<std::array<double, 6>> box(ncl);
    for (int c = 0; c < ncl; ++c) {
        for (size_t q = 0; q < leaves[c].size(); ++q) {
            const int k = leaves[c][q];
            H.cl_seg[((size_t)c * HM_SEGS + q) * 2] = seg_a[k];
            H.cl_seg[((size_t)c * HM_SEGS + q) * 2 + 1] = seg_b[k] - seg_a[k];
        }
        int m = 0;
        std::array<double, 6> b = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
        for (int k : leaves[c])
            for (int i = seg_a[k]; i < seg_b[k]; ++i) {
                H.cl_idx[(size_t)c * HM_CL + m++] = i;
                for (int d = 0; d < 3; ++d) {
                    b[d] = std::min(b[d], X[3 * (size_t)i + d]);
                    b[3 + d] = std::max(b[3 + d], X[3 * (size_t)i + d]);
                }
            }
        H.cl_size[c] = m;
        box[c] = b;
    }
    // dense by geometry: the bounding boxes touch or overlap (gap below a billionth of
    // the larger box)
    H.near.assign((size_t)ncl * ncl, 0);
    for (int I = 0; I < ncl; ++I)
        for (int J = 0; J < ncl; ++J) {
            double g2 = 0, d2 = 0;
            for (int d = 0; d < 3; ++d) {
                const double g = std::max(0.0, std::max(box[I][d] - box[J][3 + d],
                                                        box[J][d] - box[I][3 + d]));
                const double e = std::max(box[I][3 + d] - box[I][d], box[J][3 + d] - box[J][d]);
                g2 += g * g;
                d2 += e * e;
            }
            H.near[(size_t)I * ncl + J] = (I == J || g2 <= 1e-18 * d2) ? 1 : 0;
        }
}

std::string build_constraints(int n, const int *dnp, const int *dni, const double *surface,
                              const double *normals, const double *grads, const double *tmp_rhs,
                              ConstraintSet &out)
{
    std::map<int, Line> lines;
    auto add_line = [&](int r) { lines.emplace(r, Line()); };  // existing line kept
    auto add_entry = [&](int r, int c, double w) {
        Line &L = lines[r];
        for (auto &p : L.e)
            if (p.first == c) return;
        L.e.emplace_back(c, w);
    };
    for (int i = 0; i < n; ++i) {
        const int p0 = dnp[i], p1 = dnp[i + 1];
        int first = dni[p0];
        for (int p = p0; p < p1; ++p)
            if (surface[dni[p]] == 1) {
                first = dni[p];
                break;
            }
        if (i != first) continue;
        if (surface[i] == 1) {
            for (int p = p0; p < p1; ++p) {
                const int k = dni[p];
                if (k == i) continue;
                if (surface[k] == 1) {
                    if (!normals) return "Dirichlet-Dirichlet double node needs node normals";
                    const double *ni = normals + 3 * i, *nk = normals + 3 * k;
                    const double dx = nk[0] - ni[0], dy = nk[1] - ni[1], dz = nk[2] - ni[2];
                    if (std::sqrt(dx * dx + dy * dy + dz * dz) < 1e-4) {
                        add_line(k);
                        add_entry(k, i, 1.0);
                    } else {
                        if (!grads) return "sharp Dirichlet-Dirichlet edge needs surface gradients";
                        const double *gi = grads + 3 * i, *gk = grads + 3 * k;
                        const double c = ni[0] * nk[0] + ni[1] * nk[1] + ni[2] * nk[2];
                        const double gkni = gk[0] * ni[0] + gk[1] * ni[1] + gk[2] * ni[2];
                        const double gink = gi[0] * nk[0] + gi[1] * nk[1] + gi[2] * nk[2];
                        const double f = 1.0 / (1.0 - c * c);
                        add_line(k);
                        add_line(i);
                        lines[i].inhom = f * (gkni + gink * c);
                        lines[k].inhom = f * (gink + gkni * c);
                    }
                } else {
                    add_line(k);
                    lines[k].inhom = tmp_rhs[i];
                }
            }
        } else {
            for (int p = p0; p < p1; ++p) {
                const int k = dni[p];
                if (k == i) continue;
                add_line(k);
                add_entry(k, i, 1.0);
            }
        }
    }
    // close: resolve entries on constrained dofs (chains), merge, sort
    for (int pass = 0;; ++pass) {
        if (pass > 64) return "constraint chain does not terminate";
        bool changed = false;
        for (auto &kv : lines) {
            Line &L = kv.second;
            std::vector<std::pair<int, double>> ne;
            for (auto &p : L.e) {
                auto it = lines.find(p.first);
                if (it == lines.end()) {
                    ne.push_back(p);
                    continue;
                }
                if (it->first == kv.first) return "self-referencing constraint";
                changed = true;
                L.inhom += p.second * it->second.inhom;
                for (auto &q : it->second.e) ne.emplace_back(q.first, p.second * q.second);
            }
            std::sort(ne.begin(), ne.end());
            std::vector<std::pair<int, double>> merged;
            for (auto &p : ne) {
                if (!merged.empty() && merged.back().first == p.first)
                    merged.back().second += p.second;
                else
                    merged.push_back(p);
            }
            L.e.swap(merged);
        }
        if (!changed) break;
    }
    out = ConstraintSet();
    out.ptr.push_back(0);
    for (auto &kv : lines) {
        out.rows.push_back(kv.first);
        out.inhom.push_back(kv.second.inhom);
        for (auto &p : kv.second.e) {
            out.cols.push_back(p.first);
            out.w.push_back(p.second);
        }
        out.ptr.push_back((int)out.cols.size());
    }
    return "";
}

}  // namespace wb
