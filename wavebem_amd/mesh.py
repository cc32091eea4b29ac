"""Synthetic closed Q1 surface meshes: the geometry producer for the BEM hot path.

WaveBEM builds its closed towing-tank surface with OpenCASCADE inside
``NumericalTowingTank::create_initial_mesh`` (source/numerical_towing_tank.cc:582-1790)
and hands the boundary-element solver a deal.II ``DoFHandler<2,3>`` with Q1 dofs, a
``double_nodes_set`` (source/computational_domain.cc:257-307) and the Dirichlet/Neumann
masks ``surface_nodes``/``other_nodes`` (source/numerical_towing_tank.cc:1799-1807).
Neither CAD nor deal.II is available here, so this module produces the same data
directly as flat arrays:

* ``support_points``  float64 [n_dofs, 3]   collocation nodes (mapped Q1 vertices)
* ``cell_dofs``       int32   [n_cells, 4]  deal.II lexicographic vertex order
                                            v0=(0,0) v1=(1,0) v2=(0,1) v3=(1,1)
* ``cell_orient``     int8    [n_cells]     +-1, the cell direction flag: the cell normal is
                                            orient * (dx/dxi x dx/deta)/|.|, chosen so that
                                            every normal points out of the fluid
* ``cell_material``   int32   [n_cells]     patch id (FREE_SURFACE, HULL, walls ...)
* ``surface_nodes``   float64 [n_dofs]      1 on Dirichlet (free-surface) dofs
* ``other_nodes``     float64 [n_dofs]      1 - surface_nodes

Each patch owns its own vertices, so patch edges carry geometrically coincident
"double nodes", exactly the situation WaveBEM's double-node machinery handles.

The Wigley hull is y = (B/2)(1-(2x/L)^2)(1-(z/T)^2) (source/boat_surface.cc:82,93).
The tank is a closed half box (symmetry plane y=0 panelled explicitly as a Neumann
wall), so the solid-angle diagonal alpha = -H.1 (source/bem_problem.cc:609) keeps its
meaning.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

FREE_SURFACE = 1
HULL = 2
CENTREPLANE = 3
FAR_SIDE = 4
BOTTOM = 5
INFLOW = 6
OUTFLOW = 7
BOX = 8

GRAVITY = 9.81
RHO = 1025.1  # source/free_surface.cc:9417


@dataclass
class SurfaceMesh:
    support_points: np.ndarray
    cell_dofs: np.ndarray
    cell_orient: np.ndarray
    cell_material: np.ndarray
    surface_nodes: np.ndarray
    other_nodes: np.ndarray
    dof_material: np.ndarray
    info: dict = field(default_factory=dict)

    @property
    def n_dofs(self) -> int:
        return int(self.support_points.shape[0])

    @property
    def n_cells(self) -> int:
        return int(self.cell_dofs.shape[0])

    def cell_vertices(self) -> np.ndarray:
        return self.support_points[self.cell_dofs]


class _Builder:
    """Accumulates structured patches into one dof/cell list."""

    def __init__(self):
        self.points = []
        self.cells = []
        self.orient = []
        self.material = []
        self.dof_material = []
        self.n = 0

    def add_grid(self, P: np.ndarray, material: int, out_dir, keep_vertex=None, keep_cell=None):
        """Add a structured (na+1) x (nb+1) vertex grid P[a, b, 3].

        Cells (a, b) use vertices (a,b), (a+1,b), (a,b+1), (a+1,b+1) (deal.II
        lexicographic order).  ``out_dir(centre) -> vec`` gives the direction that
        the cell normal must point to (out of the fluid); the orientation flag is
        chosen from it.  ``keep_vertex``/``keep_cell`` are boolean masks for grids
        with holes.
        """
        na1, nb1, _ = P.shape
        na, nb = na1 - 1, nb1 - 1
        if keep_vertex is None:
            keep_vertex = np.ones((na1, nb1), dtype=bool)
        if keep_cell is None:
            keep_cell = np.ones((na, nb), dtype=bool)
        ids = -np.ones((na1, nb1), dtype=np.int64)
        # row-major numbering inside a patch, a fastest
        order = [(a, b) for b in range(nb1) for a in range(na1) if keep_vertex[a, b]]
        for k, (a, b) in enumerate(order):
            ids[a, b] = self.n + k
        self.points.append(np.array([P[a, b] for (a, b) in order]))
        self.dof_material.append(np.full(len(order), material, dtype=np.int32))
        self.n += len(order)
        for b in range(nb):
            for a in range(na):
                if not keep_cell[a, b]:
                    continue
                v = [ids[a, b], ids[a + 1, b], ids[a, b + 1], ids[a + 1, b + 1]]
                if min(v) < 0:
                    raise ValueError("cell uses a removed vertex")
                p0, p1, p2, p3 = P[a, b], P[a + 1, b], P[a, b + 1], P[a + 1, b + 1]
                ca = 0.5 * ((p1 - p0) + (p3 - p2))
                cb = 0.5 * ((p2 - p0) + (p3 - p1))
                nrm = np.cross(ca, cb)
                centre = 0.25 * (p0 + p1 + p2 + p3)
                s = 1 if float(np.dot(nrm, out_dir(centre))) > 0 else -1
                self.cells.append(v)
                self.orient.append(s)
                self.material.append(material)

    def finish(self) -> tuple:
        pts = np.concatenate(self.points).astype(np.float64)
        cells = np.asarray(self.cells, dtype=np.int32)
        orient = np.asarray(self.orient, dtype=np.int8)
        mat = np.asarray(self.material, dtype=np.int32)
        dmat = np.concatenate(self.dof_material)
        return pts, cells, orient, mat, dmat


def _geometric_cells(length: float, n: int, first: float) -> np.ndarray:
    """n cell sizes growing geometrically from ``first`` and summing to ``length``."""
    if n <= 0:
        return np.zeros(0)
    if first * n >= length:
        return np.full(n, length / n)
    lo, hi = 1.0, 10.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        s = first * (r ** n - 1.0) / (r - 1.0)
        if s > length:
            hi = r
        else:
            lo = r
    r = 0.5 * (lo + hi)
    sizes = first * r ** np.arange(n)
    return sizes * (length / sizes.sum())


def wigley_hull_y(x, z, L=1.0, B=0.1, T=0.0625):
    """Half-breadth of the Wigley hull (source/boat_surface.cc:82,93)."""
    x = np.asarray(x, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    return 0.5 * B * (1.0 - (2.0 * x / L) ** 2) * (1.0 - (z / T) ** 2)


def wigley_hull_normal_out_of_fluid(p, L=1.0, B=0.1, T=0.0625):
    """Unit normal of the Wigley hull pointing out of the fluid (into the hull).

    source/boat_surface.cc:111-166 (``HullNormal``) gives the normal pointing into the
    fluid for y>0; the BEM convention needs the opposite sign.
    """
    p = np.atleast_2d(p)
    x, z = p[:, 0], p[:, 2]
    nx = 4.0 * B / (L * L) * (1.0 - (z / T) ** 2) * x
    ny = np.ones_like(x)
    nz = B / (T * T) * (1.0 - (2.0 * x / L) ** 2) * z
    n = np.stack([nx, ny, nz], axis=1)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return -n


def wigley_tank(nh=40, nt=8, nu=16, nd=24, ny=20, nzd=6,
                L=1.0, B=0.1, T=0.0625, D=1.0, x0=-1.5, x1=2.5, Y=1.5,
                y_first=None) -> SurfaceMesh:
    """Closed half-box around a Wigley half hull, all patches Q1 structured grids.

    nh x nt hull cells; the free surface [x0,x1]x[0,Y] at z=0 has nu+nh+nd cells in x
    (geometrically graded away from the hull) and ny cells across; the walls use the
    same x/y nodes and nt+nzd cells in depth (nt on [-T,0], nzd graded to -D).
    """
    dxh = L / nh
    xh = np.linspace(-0.5 * L, 0.5 * L, nh + 1)
    up = _geometric_cells(-0.5 * L - x0, nu, dxh)
    dn = _geometric_cells(x1 - 0.5 * L, nd, dxh)
    xs = np.concatenate([(-0.5 * L) - np.cumsum(up)[::-1], xh, 0.5 * L + np.cumsum(dn)])
    xs[0] = x0
    xs[-1] = x1
    nx = xs.size - 1
    if y_first is None:
        y_first = dxh
    eta = np.concatenate([[0.0], np.cumsum(_geometric_cells(1.0, ny, y_first / Y))])
    eta[-1] = 1.0
    zt = np.linspace(-T, 0.0, nt + 1)
    zdeep = -T - np.concatenate([[0.0], np.cumsum(_geometric_cells(D - T, nzd, T / nt))])
    zdeep[-1] = -D
    zc = np.concatenate([zdeep[::-1][:-1], zt])  # from -D up to 0
    nzc = zc.size - 1

    b = _Builder()
    w = np.where(np.abs(xs) <= 0.5 * L, wigley_hull_y(xs, 0.0, L, B, T), 0.0)
    w[np.abs(xs) >= 0.5 * L - 1e-14] = 0.0

    # free surface z=0, normal +z
    P = np.zeros((nx + 1, ny + 1, 3))
    for a in range(nx + 1):
        P[a, :, 0] = xs[a]
        P[a, :, 1] = w[a] + eta * (Y - w[a])
    b.add_grid(P, FREE_SURFACE, lambda c: np.array([0.0, 0.0, 1.0]))

    # hull: (x, z) grid, normal into the hull (-y side)
    P = np.zeros((nh + 1, nt + 1, 3))
    for a in range(nh + 1):
        for c in range(nt + 1):
            P[a, c] = [xh[a], wigley_hull_y(xh[a], zt[c], L, B, T), zt[c]]
    P[0, :, 1] = 0.0
    P[-1, :, 1] = 0.0
    P[:, 0, 1] = 0.0
    b.add_grid(P, HULL, lambda c: np.array([0.0, -1.0, 0.0]))

    # centre plane y=0 with the hull rectangle [-L/2,L/2]x[-T,0] removed, normal -y
    P = np.zeros((nx + 1, nzc + 1, 3))
    for a in range(nx + 1):
        P[a, :, 0] = xs[a]
        P[a, :, 2] = zc
    inside_x = np.abs(xs) < 0.5 * L - 1e-12
    inside_z = zc > -T + 1e-12
    keep_v = ~(inside_x[:, None] & inside_z[None, :])
    xc = 0.5 * (xs[1:] + xs[:-1])
    zcc = 0.5 * (zc[1:] + zc[:-1])
    keep_c = ~((np.abs(xc) < 0.5 * L)[:, None] & (zcc > -T)[None, :])
    b.add_grid(P, CENTREPLANE, lambda c: np.array([0.0, -1.0, 0.0]), keep_v, keep_c)

    # far side y=Y, normal +y
    P = np.zeros((nx + 1, nzc + 1, 3))
    for a in range(nx + 1):
        P[a, :, 0] = xs[a]
        P[a, :, 1] = Y
        P[a, :, 2] = zc
    b.add_grid(P, FAR_SIDE, lambda c: np.array([0.0, 1.0, 0.0]))

    # bottom z=-D, normal -z
    P = np.zeros((nx + 1, ny + 1, 3))
    for a in range(nx + 1):
        P[a, :, 0] = xs[a]
        P[a, :, 1] = eta * Y
        P[a, :, 2] = -D
    b.add_grid(P, BOTTOM, lambda c: np.array([0.0, 0.0, -1.0]))

    # inflow x=x0 normal -x, outflow x=x1 normal +x
    for xv, mat, sgn in ((x0, INFLOW, -1.0), (x1, OUTFLOW, 1.0)):
        P = np.zeros((ny + 1, nzc + 1, 3))
        for a in range(ny + 1):
            P[a, :, 0] = xv
            P[a, :, 1] = eta[a] * Y
            P[a, :, 2] = zc
        b.add_grid(P, mat, (lambda s: (lambda c: np.array([s, 0.0, 0.0])))(sgn))

    pts, cells, orient, mat, dmat = b.finish()
    surf = (dmat == FREE_SURFACE).astype(np.float64)
    info = dict(kind="wigley_tank", L=L, B=B, T=T, D=D, x0=x0, x1=x1, Y=Y,
                nh=nh, nt=nt, nu=nu, nd=nd, ny=ny, nzd=nzd,
                n_per_patch={int(m): int((dmat == m).sum()) for m in np.unique(dmat)})
    return SurfaceMesh(pts, cells, orient, mat, surf, 1.0 - surf, dmat, info)


# approximate dof targets -> generator parameters (counts reported by the mesh itself)
WIGLEY_PRESETS = {
    "2.5k": dict(nh=24, nt=6, nu=10, nd=14, ny=14, nzd=4),
    "10k": dict(nh=48, nt=12, nu=20, nd=28, ny=30, nzd=6),
    "20k": dict(nh=70, nt=16, nu=28, nd=40, ny=42, nzd=8),
    "40k": dict(nh=100, nt=22, nu=40, nd=56, ny=60, nzd=10),
    "60k": dict(nh=120, nt=28, nu=50, nd=70, ny=72, nzd=12),
    "70k": dict(nh=130, nt=30, nu=54, nd=76, ny=78, nzd=12),
}


def wigley_preset(name: str) -> SurfaceMesh:
    return wigley_tank(**WIGLEY_PRESETS[name])


def box_mesh(n=4, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), dirichlet_top=False,
             dirichlet_faces=(), split_top=False) -> SurfaceMesh:
    """Closed axis-aligned box, six separately numbered faces, normals outward.

    Used for the solid-angle checks: alpha = 1/2 on faces, 1/4 on edges, 1/8 at
    corners for the interior problem.  Face (axis, side) has material
    BOX + 2 axis + side; ``dirichlet_faces`` lists the (axis, side) faces whose dofs are
    Dirichlet (``dirichlet_top`` adds (2, 1)).  Two adjacent Dirichlet faces give sharp
    Dirichlet-Dirichlet double nodes; ``split_top`` builds the top face as two patches
    with their own vertices, whose shared line (x = midpoint) carries flat ones.
    """
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    b = _Builder()
    t = np.linspace(0.0, 1.0, n + 1)
    centre = 0.5 * (lo + hi)
    for axis in range(3):
        for side in (0, 1):
            u, v = [k for k in range(3) if k != axis]
            P = np.zeros((n + 1, n + 1, 3))
            for a in range(n + 1):
                for c in range(n + 1):
                    p = np.empty(3)
                    p[axis] = hi[axis] if side else lo[axis]
                    p[u] = lo[u] + t[a] * (hi[u] - lo[u])
                    p[v] = lo[v] + t[c] * (hi[v] - lo[v])
                    P[a, c] = p
            mat = BOX + 2 * axis + side
            if split_top and (axis, side) == (2, 1):
                h = n // 2
                b.add_grid(P[:h + 1], mat, lambda c, cc=centre: c - cc)
                b.add_grid(P[h:], mat, lambda c, cc=centre: c - cc)
            else:
                b.add_grid(P, mat, lambda c, cc=centre: c - cc)
    pts, cells, orient, mat, dmat = b.finish()
    faces = set(tuple(f) for f in dirichlet_faces)
    if dirichlet_top:
        faces.add((2, 1))
    surf = np.zeros(pts.shape[0])
    for axis, side in faces:
        surf[dmat == BOX + 2 * axis + side] = 1.0
    return SurfaceMesh(pts, cells, orient, mat, surf, 1.0 - surf, dmat,
                       dict(kind="box", n=n))


def cube_sphere_mesh(n=4, radius=1.0, dirichlet_upper=True) -> SurfaceMesh:
    """Sphere as six projected cube faces (double nodes on the cube edges)."""
    b = _Builder()
    t = np.linspace(-1.0, 1.0, n + 1)
    for axis in range(3):
        for side in (-1.0, 1.0):
            u, v = [k for k in range(3) if k != axis]
            P = np.zeros((n + 1, n + 1, 3))
            for a in range(n + 1):
                for c in range(n + 1):
                    p = np.empty(3)
                    p[axis] = side
                    p[u] = math.tan(0.25 * math.pi * t[a])
                    p[v] = math.tan(0.25 * math.pi * t[c])
                    P[a, c] = radius * p / np.linalg.norm(p)
            b.add_grid(P, BOX + axis, lambda c: c)
    pts, cells, orient, mat, dmat = b.finish()
    if dirichlet_upper:
        surf = (pts[:, 2] > 1e-12).astype(np.float64)
    else:
        surf = np.zeros(pts.shape[0])
    return SurfaceMesh(pts, cells, orient, mat, surf, 1.0 - surf, dmat,
                       dict(kind="cube_sphere", n=n, radius=radius))


def tri_block_mesh(n=4, curvature=0.2, centre=(0.0, 0.0, 0.0), radius=1.0) -> SurfaceMesh:
    """An open curved patch that is not a structured grid: a triangle cut into three
    quadrilateral blocks (n x n cells each) that share their edges, so the centre vertex
    has three cells around it.  z = curvature (x^2 + y^2) over the plane; all dofs are
    Dirichlet.  Used to exercise the column-tile assembly of unstructured parts."""
    c = np.asarray(centre, dtype=np.float64)
    ang = np.pi / 2 + np.array([0.0, 2 * np.pi / 3, 4 * np.pi / 3])
    V = [np.array([radius * np.cos(a), radius * np.sin(a)]) for a in ang]
    O = np.zeros(2)
    mid = [(V[k] + V[(k + 1) % 3]) / 2 for k in range(3)]
    blocks = [(V[0], mid[0], mid[2], O), (V[1], mid[1], mid[0], O), (V[2], mid[2], mid[1], O)]
    t = np.linspace(0.0, 1.0, n + 1)
    keyed, pts, cells = {}, [], []

    def dof(p):
        k = (round(p[0], 12), round(p[1], 12))
        if k not in keyed:
            keyed[k] = len(pts)
            pts.append(np.array([p[0], p[1], curvature * (p[0] ** 2 + p[1] ** 2)]) + c)
        return keyed[k]

    for (p00, p10, p01, p11) in blocks:
        ids = [[dof((1 - a) * (1 - b) * p00 + a * (1 - b) * p10 + (1 - a) * b * p01 + a * b * p11)
                for b in t] for a in t]
        for j in range(n):
            for i in range(n):
                cells.append([ids[i][j], ids[i + 1][j], ids[i][j + 1], ids[i + 1][j + 1]])
    pts = np.array(pts)
    cells = np.asarray(cells, dtype=np.int32)
    orient = np.empty(len(cells), dtype=np.int8)
    for k, cv in enumerate(cells):
        p0, p1, p2, p3 = pts[cv]
        nrm = np.cross(0.5 * ((p1 - p0) + (p3 - p2)), 0.5 * ((p2 - p0) + (p3 - p1)))
        orient[k] = 1 if nrm[2] > 0 else -1
    mat = np.full(len(cells), BOX, dtype=np.int32)
    dmat = np.full(len(pts), BOX, dtype=np.int32)
    surf = np.ones(len(pts))
    return SurfaceMesh(pts, cells, orient, mat, surf, 1.0 - surf, dmat,
                       dict(kind="tri_block", n=n))


def merge_meshes(*meshes: SurfaceMesh) -> SurfaceMesh:
    """Disjoint union of meshes (dofs and cells renumbered consecutively)."""
    off, cells = 0, []
    for m in meshes:
        cells.append(m.cell_dofs + off)
        off += m.n_dofs
    cat = lambda f: np.concatenate([getattr(m, f) for m in meshes])  # noqa: E731
    return SurfaceMesh(cat("support_points"), np.concatenate(cells).astype(np.int32),
                       cat("cell_orient"), cat("cell_material"), cat("surface_nodes"),
                       cat("other_nodes"), cat("dof_material"), dict(kind="merged"))


def double_nodes_csr(points: np.ndarray, tol: float = 1e-8):
    """Sets of geometrically coincident dofs (each set contains the dof itself).

    Same semantics as ComputationalDomain::generate_double_nodes_set
    (source/computational_domain.cc:257-307: distance < 1e-8), computed with a
    sort instead of the O(n^2) double loop.  Returns (ptr, idx) CSR, sorted indices.
    """
    n = points.shape[0]
    q = np.floor(points / (4 * tol)).astype(np.int64)
    order = np.lexsort((q[:, 2], q[:, 1], q[:, 0]))
    sets = [[i] for i in range(n)]
    # neighbours can only be within the same or adjacent grid cells; check runs
    keyed = {}
    for i in order:
        keyed.setdefault(tuple(q[i]), []).append(int(i))
    for i in range(n):
        qi = q[i]
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in keyed.get((qi[0] + dx, qi[1] + dy, qi[2] + dz), ()):
                        if j != i and np.linalg.norm(points[i] - points[j]) < tol:
                            sets[i].append(j)
    ptr = np.zeros(n + 1, dtype=np.int32)
    idx = []
    for i in range(n):
        s = sorted(set(sets[i]))
        idx.extend(s)
        ptr[i + 1] = ptr[i] + len(s)
    return ptr, np.asarray(idx, dtype=np.int32)


def wigley_bc(mesh: SurfaceMesh, froude=0.3, seed=0, phi_amp=1e-3):
    """Boundary data vector ``tmp_rhs`` in WaveBEM's convention.

    Dirichlet (free-surface) slots hold phi ~ U[-phi_amp, phi_amp] (numpy PCG64,
    ``seed``); hull slots hold dphi/dn = -n . V_inf with V_inf = (U,0,0) and
    U = Fr sqrt(g L) (source/free_surface.cc:8833); other walls hold 0
    (source/free_surface.cc:8845-8849).
    """
    L = mesh.info.get("L", 1.0)
    B = mesh.info.get("B", 0.1)
    T = mesh.info.get("T", 0.0625)
    U = froude * math.sqrt(GRAVITY * L)
    rng = np.random.Generator(np.random.PCG64(seed))
    bc = np.zeros(mesh.n_dofs)
    fs = mesh.dof_material == FREE_SURFACE
    bc[fs] = rng.uniform(-phi_amp, phi_amp, size=int(fs.sum()))
    hull = mesh.dof_material == HULL
    n = wigley_hull_normal_out_of_fluid(mesh.support_points[hull], L, B, T)
    bc[hull] = -U * n[:, 0]
    return bc, U
