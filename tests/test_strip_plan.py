"""The strip schedule of the regular-cell assembly, checked on the CPU by replaying the
strip kernel's slot logic (k_assemble_strips, wb_assemble.hip) over the plan the library
builds: every (cell, vertex) incidence of a strip column must reach that column's slot
exactly once, in the line the column belongs to, and every strip column must be written
exactly once.  The kernel adds the cell values it is given to these slots, so a plan that
passes produces each matrix entry as the sum over the cells incident to its column."""
import numpy as np
import pytest

from tests import helpers as T
from wavebem_amd import bem as B
from wavebem_amd import mesh as M

W = 8  # STRIP_W


def replay(mesh, plan):
    """Follows the kernel gap by gap: the cell at position k of gap t adds its canonical
    vertices 0, 1 to positions k - 1, k of line t (lo) and 2, 3 to those of line t + 1
    (up); after the gap, lo is written (the positions in its mask, at consecutive columns
    from its first column) and up becomes lo."""
    n = mesh.n_dofs
    perm = plan["perm"]
    cis = plan["cell_int_s"]
    own = perm[mesh.cell_dofs]
    assert np.array_equal(np.sort(cis, axis=1), np.sort(own, axis=1))
    written = np.zeros(n, dtype=np.int64)
    reached = np.zeros(n, dtype=np.int64)
    for g0, ng in plan["tasks"]:
        cols = set()
        for g in range(g0, g0 + ng):
            col0, mask = plan["gaps"][g][1], plan["gaps"][g][2]
            cols.update(range(col0, col0 + bin(mask).count("1")))
        lo = [[] for _ in range(W)]
        up = [[] for _ in range(W)]
        for g in range(g0, g0 + ng):
            cmask, col0, lmask, _ = plan["gaps"][g]
            if g == g0:
                assert lmask == 0  # gap -1 completes the line before the task
            for k in range(W + 1):
                cell = plan["gap_cells"][g][k]
                assert (cell >= 0) == bool(cmask >> k & 1)
                if cell < 0:
                    continue
                v = cis[cell]
                if k >= 1:
                    lo[k - 1].append(v[0])
                    up[k - 1].append(v[2])
                if k < W:
                    lo[k].append(v[1])
                    up[k].append(v[3])
            for p in range(W):
                if lmask >> p & 1:
                    col = col0 + bin(lmask & ((1 << p) - 1)).count("1")
                    written[col] += 1
                    assert all(v == col for v in lo[p]), (lo[p], col)
                    reached[col] += len(lo[p])
                else:
                    # spare slot or a line outside the task: only other tasks' columns
                    assert all(v not in cols for v in lo[p])
            lo, up = up, [[] for _ in range(W)]
        # the line after the task (the last up) belongs to the next task
        assert all(v not in cols for slot in lo for v in slot)
    incid = np.bincount(own.reshape(-1), minlength=n)
    sd = plan["strip_dofs"]
    assert np.all(written[:sd] == 1)
    assert np.all(written[sd:] == 0)
    assert np.array_equal(reached[:sd], incid[:sd])
    return sd


@pytest.mark.parametrize("kind", ["wigley_small", "2.5k", "10k", "box", "sphere",
                                  "tri_block", "mixed"])
def test_strip_schedule_reaches_every_incidence_once(kind):
    if kind == "wigley_small":
        m = T.wigley_small()
    elif kind in ("2.5k", "10k"):
        m = T.wigley_preset(kind)
    elif kind == "box":
        m = M.box_mesh(n=6)
    elif kind == "sphere":
        m = M.cube_sphere_mesh(n=9)
    elif kind == "tri_block":
        m = M.tri_block_mesh(n=5)
    else:
        m = M.merge_meshes(M.box_mesh(n=5), M.tri_block_mesh(n=4, centre=(3.0, 0, 0)))
    plan = B.host_strip_plan(m)
    sd = replay(m, plan)
    if kind == "tri_block":
        # three blocks around a vertex of valence 3: no structured grid, column tiles only
        assert sd == 0 and plan["n_tiles"] > 0
    elif kind == "mixed":
        assert 0 < sd < m.n_dofs and plan["n_tiles"] > 0
    else:
        assert sd == m.n_dofs and plan["n_tiles"] == 0
    # the device order is a permutation, with the strip columns first
    assert np.array_equal(np.sort(plan["perm"]), np.arange(m.n_dofs))


def test_strip_schedule_evaluates_cells_about_once():
    m = T.wigley_preset("10k")
    plan = B.host_strip_plan(m)
    visits = plan["cells"] / m.n_cells
    assert 1.0 <= visits < 1.15
    assert np.all(np.array([bin(x).count("1") for x in plan["gaps"][:, 2]]) <= W)


def test_tiles_only_plan_has_no_strips():
    m = T.wigley_small()
    plan = B.host_strip_plan(m, strips=0)
    assert plan["strip_dofs"] == 0 and len(plan["tasks"]) == 0 and plan["n_tiles"] > 0
